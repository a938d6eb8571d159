// Bootstrap outputs in the reference's layouts (io/bootstrap_file.cpp:17-185);
// the layouts are described in io/bootstrap_file.hpp.
#include "io/bootstrap_file.hpp"

#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace ioh {

namespace {
constexpr size_t BLOCK = 2880;

struct Card {
    std::string key, value;   // value already formatted; empty key: COMMENT
};

std::string fmt_double(double v) {
    char b[32];
    std::snprintf(b, sizeof b, "%.17G", v);
    std::string s(b);
    if (s.find_first_of(".EN") == std::string::npos) s += ".";   // a FITS real, not an integer
    return s;
}
std::string fmt_string(const std::string &v) {
    std::string s = "'" + v;
    while (s.size() < 9) s += ' ';
    return s + "'";
}

void put(std::vector<char> &out, const std::string &card) {
    std::string c = card.substr(0, 80);
    c.resize(80, ' ');
    out.insert(out.end(), c.begin(), c.end());
}

// one HDU: the structural keys, then cards; data as big-endian doubles
void write_hdu(std::FILE *f, bool primary, const std::vector<long> &naxes, const std::vector<Card> &cards,
               const double *data) {
    std::vector<char> h;
    char b[96];
    if (primary) {
        put(h, "SIMPLE  =                    T");
    } else {
        std::snprintf(b, sizeof b, "XTENSION= %-20s", "'IMAGE   '");
        put(h, b);
    }
    put(h, "BITPIX  =                  -64");
    std::snprintf(b, sizeof b, "NAXIS   = %20zu", naxes.size());
    put(h, b);
    size_t n = naxes.empty() ? 0 : 1;
    for (size_t a = 0; a < naxes.size(); ++a) {
        std::snprintf(b, sizeof b, "NAXIS%-3zu= %20ld", a + 1, naxes[a]);
        put(h, b);
        n *= (size_t)naxes[a];
    }
    if (primary) {
        put(h, "EXTEND  =                    T");
    } else {
        put(h, "PCOUNT  =                    0");
        put(h, "GCOUNT  =                    1");
    }
    for (const Card &c : cards) {
        if (c.key.empty()) {
            put(h, "COMMENT " + c.value);
        } else {
            std::snprintf(b, sizeof b, "%-8s= %20s", c.key.c_str(), c.value.c_str());
            put(h, b);
        }
    }
    put(h, "END");
    h.resize((h.size() + BLOCK - 1) / BLOCK * BLOCK, ' ');
    std::vector<unsigned char> d((n * 8 + BLOCK - 1) / BLOCK * BLOCK, 0);
    for (size_t i = 0; i < n; ++i) {
        unsigned char v[8];
        std::memcpy(v, data + i, 8);
        for (int k = 0; k < 8; ++k) d[i * 8 + k] = v[7 - k];
    }
    if (std::fwrite(h.data(), 1, h.size(), f) != h.size() || std::fwrite(d.data(), 1, d.size(), f) != d.size())
        throw std::runtime_error("FITS ERROR writing");
}

std::FILE *create(const std::string &path) {
    std::FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("FITS ERROR could not create " + path);
    return f;
}

// reads the image HDU named extname (EXTNAME) from a file this code wrote
void read_image(const std::vector<unsigned char> &buf, const std::string &extname,
                std::map<std::string, std::string> &keys, std::vector<double> &values) {
    size_t off = 0;
    while (off + BLOCK <= buf.size()) {
        keys.clear();
        bool end = false;
        while (!end) {
            if (off + BLOCK > buf.size()) throw std::runtime_error("FITS ERROR truncated header");
            for (size_t c = 0; c < BLOCK && !end; c += 80) {
                const std::string card((const char *)buf.data() + off + c, 80);
                std::string key = card.substr(0, 8);
                key.erase(key.find_last_not_of(' ') + 1);
                if (key == "END") end = true;
                else if (card.compare(8, 2, "= ") == 0) {
                    std::string v = card.substr(10);
                    v.erase(0, v.find_first_not_of(' '));
                    if (!v.empty() && v[0] == '\'') v = v.substr(1, v.find('\'', 1) - 1);
                    v.erase(v.find_last_not_of(' ') + 1);
                    keys[key] = v;
                }
            }
            off += BLOCK;
        }
        const int naxis = std::stoi(keys.at("NAXIS"));
        size_t n = naxis ? 1 : 0;
        for (int a = 1; a <= naxis; ++a) n *= std::stoul(keys.at("NAXIS" + std::to_string(a)));
        if (keys.count("EXTNAME") && keys["EXTNAME"] == extname) {
            if (off + n * 8 > buf.size()) throw std::runtime_error("FITS ERROR truncated data");
            values.resize(n);
            for (size_t i = 0; i < n; ++i) {
                unsigned char v[8];
                for (int k = 0; k < 8; ++k) v[k] = buf[off + i * 8 + 7 - k];
                std::memcpy(&values[i], v, 8);
            }
            return;
        }
        off += (n * 8 + BLOCK - 1) / BLOCK * BLOCK;
    }
    throw std::runtime_error("FITS ERROR HDU " + extname + " not found");
}
}  // namespace

// io/bootstrap_file.cpp:17-78
std::string saveBootstrapRealizations(const std::string &base, const double *allpowers, const double *invfisher,
                                      unsigned int nboots, int nk, int nz, bool fastbootstrap,
                                      const char *comment) {
    const std::string out_fname = "!" + base + "-bootstrap-realizations.fits";
    std::FILE *f = create(out_fname.substr(1));
    const long nkz = (long)nk * nz;
    try {
        std::vector<Card> c = {{"EXTNAME", fmt_string("REALIZATIONS")},
                               {"NBOOTS", std::to_string(nboots)},
                               {"NK", std::to_string(nk)},
                               {"NZ", std::to_string(nz)},
                               {"FASTBOOT", fastbootstrap ? "T" : "F"}};
        if (comment) c.push_back({"", comment});
        write_hdu(f, true, {nkz, (long)nboots}, c, allpowers);
        std::vector<Card> c2 = {{"EXTNAME", fmt_string("SOLV_INVF")},
                                {"NK", std::to_string(nk)},
                                {"NZ", std::to_string(nz)},
                                {"FASTBOOT", fastbootstrap ? "T" : "F"},
                                {"", "Inverse Fisher solver matrix. Not necessarily equal to the inverse Fisher "
                                     "or symmetric."}};
        if (comment) c2.push_back({"", comment});
        write_hdu(f, false, {nkz, nkz}, c2, invfisher);
    } catch (...) {
        std::fclose(f);
        throw;
    }
    if (std::fclose(f) != 0) throw std::runtime_error("FITS ERROR closing " + out_fname.substr(1));
    return out_fname;
}

// io/bootstrap_file.cpp:81-120
void readBootstrapRealizations(const std::string &fname, std::unique_ptr<double[]> &allpowers,
                               std::unique_ptr<double[]> &invfisher, unsigned int &nboots, int &nk, int &nz,
                               bool &fastbootstrap) {
    std::FILE *f = std::fopen(fname.c_str(), "rb");
    if (!f) throw std::runtime_error("FITS ERROR could not open " + fname);
    std::vector<unsigned char> buf;
    unsigned char tmp[1 << 16];
    size_t r;
    while ((r = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + r);
    std::fclose(f);
    std::map<std::string, std::string> keys;
    std::vector<double> v;
    read_image(buf, "REALIZATIONS", keys, v);
    nboots = (unsigned)std::stoul(keys.at("NBOOTS"));
    nk = std::stoi(keys.at("NK"));
    nz = std::stoi(keys.at("NZ"));
    fastbootstrap = keys.at("FASTBOOT") == "T";
    allpowers = std::make_unique<double[]>(v.size());
    std::copy(v.begin(), v.end(), allpowers.get());
    read_image(buf, "SOLV_INVF", keys, v);
    invfisher = std::make_unique<double[]>(v.size());
    std::copy(v.begin(), v.end(), invfisher.get());
}

// io/bootstrap_file.cpp:122-131
BootstrapChunksFile::BootstrapChunksFile(const std::string &base, int thispe) {
    file = create(base + "-bootchunks-" + std::to_string(thispe) + ".fits");
}

BootstrapChunksFile::~BootstrapChunksFile() {
    if (!file) return;
    if (!any) {   // cfitsio leaves an empty primary HDU in a file with no chunks
        try {
            write_hdu(file, true, {}, {}, nullptr);
        } catch (...) {
        }
    }
    std::fclose(file);
}

// io/bootstrap_file.cpp:134-185
void BootstrapChunksFile::writeChunk(const double *pk, const double *nk, const double *tk, const double *fisher,
                                     int ndim, int fisher_index_start, long id, double z_qso, double ra,
                                     double dec) {
    const long cf_size = 3L * ndim + ((long)ndim * (ndim + 1)) / 2;
    std::vector<double> data(cf_size);
    std::copy(pk, pk + ndim, data.begin());
    std::copy(nk, nk + ndim, data.begin() + ndim);
    std::copy(tk, tk + ndim, data.begin() + 2 * ndim);
    double *v = data.data() + 3 * ndim;
    for (int d = 0; d < ndim; ++d)   // mxhelp::getDiagonal(fisher, ndim, d, v): F[i][i + d]
        for (int i = 0; i + d < ndim; ++i) *v++ = fisher[(size_t)i * ndim + i + d];
    const std::vector<Card> c = {{"TARGETID", std::to_string(id)},
                                 {"ZQSO", fmt_double(z_qso)},
                                 {"RA", fmt_double(ra)},
                                 {"DEC", fmt_double(dec)},
                                 {"NQDIM", std::to_string(ndim)},
                                 {"ISTART", std::to_string(fisher_index_start)}};
    write_hdu(file, !any, {cf_size}, c, data.data());
    any = true;
}

}  // namespace ioh
