// Reader of picca delta files (one binary-table HDU per spectrum), the
// counterpart of qio::PiccaFile (io/qso_file.cpp:376-677) for the drop-in's
// qio::QSOFile: "file.fits[.gz][N]" names HDU N + 1 (HDU 2 without a
// bracket, decomposeFname, :376-386); the header gives NAXIS2 (pixels),
// TARGETID or THING_ID, Z, RA, DEC, MEANRESO (R in km/s), MEANSNR, DLL
// (dv = DLL c ln 10) or DLAMBDA if present, and EXPID, NIGHT, FIBER,
// PETAL_LOC if present (:538-600); the columns LOGLAM (10^x) or LAMBDA,
// DELTA or DELTA_BLIND, IVAR (:612-655), and RESOMAT, whose rows of ndiags
// values (TDIM) become the diagonal storage ndiags x N (:657-677, with
// DiaMatrix::orderTranspose).  R_fwhm = int(c / R_kms / 2.35482... / 100 +
// 0.5) * 100 (:91-101).  Compressed files are inflated with zlib.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "io/qso_file.hpp"

namespace qio {

namespace {
constexpr size_t BLOCK = 2880;

std::vector<unsigned char> read_file(const std::string &path) {
    gzFile f = gzopen(path.c_str(), "rb");   // reads plain files as they are
    if (!f) throw std::runtime_error("Cannot open file: " + path);
    std::vector<unsigned char> buf;
    unsigned char tmp[1 << 16];
    int n;
    while ((n = gzread(f, tmp, sizeof tmp)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    const bool bad = n < 0;
    gzclose(f);
    if (bad) throw std::runtime_error("FITS ERROR reading " + path);
    return buf;
}

struct Header {
    std::map<std::string, std::string> kv;
    bool has(const std::string &k) const { return kv.count(k) > 0; }
    const std::string &at(const std::string &k) const {
        auto it = kv.find(k);
        if (it == kv.end()) throw std::runtime_error("FITS ERROR keyword " + k + " not found");
        return it->second;
    }
    double d(const std::string &k) const {
        std::string v = at(k);
        std::replace(v.begin(), v.end(), 'D', 'E');   // Fortran exponents
        return std::stod(v);
    }
    long long i(const std::string &k) const { return std::stoll(at(k)); }
};

// parses the header at off; returns the offset of its data
size_t parse_header(const std::vector<unsigned char> &buf, size_t off, Header &h) {
    for (;;) {
        if (off + BLOCK > buf.size()) throw std::runtime_error("FITS ERROR truncated header");
        for (size_t c = 0; c < BLOCK; c += 80) {
            const std::string card((const char *)buf.data() + off + c, 80);
            std::string key = card.substr(0, 8);
            key.erase(key.find_last_not_of(' ') + 1);
            if (key == "END") return off + BLOCK;
            size_t vpos = 10;
            if (key == "HIERARCH") {   // keywords longer than 8 characters (PETAL_LOC)
                const size_t eq = card.find('=', 9);
                if (eq == std::string::npos) continue;
                key = card.substr(9, eq - 9);
                key.erase(0, key.find_first_not_of(' '));
                key.erase(key.find_last_not_of(' ') + 1);
                vpos = eq + 1;
            } else if (card.compare(8, 2, "= ") != 0) {
                continue;
            }
            std::string val = card.substr(vpos);
            val.erase(0, val.find_first_not_of(' '));
            if (!val.empty() && val[0] == '\'') {
                const size_t q = val.find('\'', 1);
                val = val.substr(1, q == std::string::npos ? std::string::npos : q - 1);
                val.erase(val.find_last_not_of(' ') + 1);
            } else {
                const size_t slash = val.find('/');
                if (slash != std::string::npos) val = val.substr(0, slash);
                val.erase(val.find_last_not_of(' ') + 1);
            }
            h.kv[key] = val;
        }
        off += BLOCK;
    }
}

// the HDU's data size in bytes; a size beyond `limit` (the bytes left in
// the file), a negative axis or an overflow is a malformed file
size_t data_bytes(const Header &h, size_t limit) {
    auto bad = [] { return std::runtime_error("FITS ERROR malformed data size"); };
    const long long naxis = h.has("NAXIS") ? h.i("NAXIS") : 0;
    if (naxis < 0 || naxis > 999) throw bad();
    const long long bitpix = h.has("BITPIX") ? std::llabs(h.i("BITPIX")) : 8;
    const long long pcount = h.has("PCOUNT") ? h.i("PCOUNT") : 0;
    if (bitpix % 8 != 0 || bitpix == 0 || bitpix > 64 || pcount < 0) throw bad();
    size_t n = naxis > 0 ? (size_t)bitpix / 8 : 0;
    for (long long a = 1; a <= naxis; ++a) {
        const long long d = h.i("NAXIS" + std::to_string(a));
        if (d < 0) throw bad();
        if (d > 0 && n > limit / (size_t)d) throw bad();
        n *= (size_t)d;
    }
    if ((size_t)pcount > limit - std::min(n, limit)) throw bad();
    return n + (size_t)pcount;
}

struct Column {
    std::string name;
    char type = 0;     // D, E, J, K, I
    long long repeat = 1;
    size_t offset = 0;  // byte offset in the row
    long long tdim0 = -1;
};

size_t type_bytes(char t) {
    switch (t) {
    case 'D': case 'K': return 8;
    case 'E': case 'J': return 4;
    case 'I': return 2;
    case 'B': case 'L': case 'A': return 1;
    default: throw std::runtime_error(std::string("FITS ERROR unsupported column type ") + t);
    }
}

double get_value(const unsigned char *p, char t) {
    unsigned char b[8];
    const size_t n = type_bytes(t);
    for (size_t k = 0; k < n; ++k) b[k] = p[n - 1 - k];   // big-endian
    switch (t) {
    case 'D': { double v; std::memcpy(&v, b, 8); return v; }
    case 'E': { float v; std::memcpy(&v, b, 4); return v; }
    case 'K': { long long v; std::memcpy(&v, b, 8); return (double)v; }
    case 'J': { int v; std::memcpy(&v, b, 4); return v; }
    case 'I': { short v; std::memcpy(&v, b, 2); return v; }
    default: return b[0];
    }
}

std::string upper(std::string s) {
    for (char &ch : s) ch = (char)std::toupper((unsigned char)ch);
    return s;
}
}  // namespace

void readPiccaFile(const std::string &fname_qso, QSOFile &q) {
    // decomposeFname: "base[n]" -> HDU n + 1, else HDU 2
    std::string base = fname_qso;
    int hdunum = 2;
    const size_t i1 = fname_qso.rfind('['), i2 = fname_qso.rfind(']');
    if (i1 != std::string::npos && i2 != std::string::npos && i2 > i1) {
        base = fname_qso.substr(0, i1);
        hdunum = std::stoi(fname_qso.substr(i1 + 1, i2 - i1 - 1)) + 1;
    }
    const std::vector<unsigned char> buf = read_file(base);
    size_t off = 0;
    Header h;
    for (int k = 1;; ++k) {
        Header hk;
        const size_t data = parse_header(buf, off, hk);
        if (k == hdunum) {
            h = hk;
            off = data;
            break;
        }
        off = data + (data_bytes(hk, buf.size() - data) + BLOCK - 1) / BLOCK * BLOCK;
        if (off >= buf.size()) throw std::runtime_error("FITS ERROR HDU " + std::to_string(hdunum) + " not found");
    }
    if (!h.has("XTENSION") || h.at("XTENSION") != "BINTABLE") throw std::runtime_error("HDU type is not BINARY!");
    const long long rowb_ll = h.i("NAXIS1"), N = h.i("NAXIS2"), nf_ll = h.i("TFIELDS");
    if (rowb_ll <= 0 || N < 0 || nf_ll <= 0 || nf_ll > 999) throw std::runtime_error("FITS ERROR malformed table header");
    const size_t rowb = (size_t)rowb_ll;
    const int nf = (int)nf_ll;
    std::vector<Column> cols(nf);
    size_t pos = 0;
    for (int c = 0; c < nf; ++c) {
        const std::string k = std::to_string(c + 1);
        Column &col = cols[c];
        col.name = upper(h.at("TTYPE" + k));
        const std::string form = h.at("TFORM" + k);
        size_t p = 0;
        while (p < form.size() && std::isdigit((unsigned char)form[p])) ++p;
        col.repeat = p ? std::stoll(form.substr(0, p)) : 1;
        if (col.repeat < 0 || (size_t)col.repeat > rowb) throw std::runtime_error("FITS ERROR malformed TFORM" + k);
        col.type = (char)std::toupper((unsigned char)form[p]);
        col.offset = pos;
        pos += (size_t)col.repeat * type_bytes(col.type);
        if (h.has("TDIM" + k)) {   // "(ndiags,N)": the first axis is the row's values
            const std::string td = h.at("TDIM" + k);
            col.tdim0 = std::stoll(td.substr(td.find('(') + 1));
        }
    }
    if (pos != rowb) throw std::runtime_error("FITS ERROR row width does not match the columns");
    if ((size_t)N > (buf.size() - off) / rowb) throw std::runtime_error("FITS ERROR truncated table");
    auto find = [&](const char *name) -> const Column * {
        for (const Column &c : cols)
            if (c.name == name) return &c;
        return nullptr;
    };
    auto read_col = [&](const Column &c, long long nper) {
        std::vector<double> out((size_t)N * nper);
        const size_t tb = type_bytes(c.type);
        for (long long r = 0; r < N; ++r)
            for (long long e = 0; e < nper; ++e)
                out[(size_t)r * nper + e] = get_value(buf.data() + off + r * rowb + c.offset + e * tb, c.type);
        return out;
    };
    // readParameters (io/qso_file.cpp:538-600)
    if (h.has("TARGETID")) q.id = h.i("TARGETID");
    else if (h.has("THING_ID")) q.id = h.i("THING_ID");
    else throw std::runtime_error("Header must have TARGETID  or THING_ID!");
    q.z_qso = h.d("Z");
    q.ra = h.d("RA");
    q.dec = h.d("DEC");
    q.R_kms = h.d("MEANRESO");
    q.snr = h.d("MEANSNR");
    const double LN10 = 2.30258509299;
    q.dv_kms = h.has("DLL") ? h.d("DLL") * SPEED_OF_LIGHT * LN10 : -1;
    q.dlambda = h.has("DLAMBDA") ? h.d("DLAMBDA") : -1;
    q.expid = h.has("EXPID") ? (int)h.i("EXPID") : -1;
    q.night = h.has("NIGHT") ? (int)h.i("NIGHT") : -1;
    q.fiber = h.has("FIBER") ? (int)h.i("FIBER") : -1;
    q.petal = h.has("PETAL_LOC") ? (int)h.i("PETAL_LOC") : -1;
    q.R_fwhm = int(SPEED_OF_LIGHT / q.R_kms / ONE_SIGMA_2_FWHM / 100 + 0.5) * 100;
    // readData (io/qso_file.cpp:612-655)
    std::vector<double> wave, delta, ivar;
    if (const Column *c = find("LOGLAM")) {
        wave = read_col(*c, 1);
        for (double &l : wave) l = std::pow(10, l);
    } else if (const Column *c = find("LAMBDA")) {
        wave = read_col(*c, 1);
    } else {
        throw std::runtime_error("Wavelength must be in LOGLAM or LAMBDA!");
    }
    const Column *cd = find("DELTA");
    if (!cd) cd = find("DELTA_BLIND");
    if (!cd) throw std::runtime_error("Deltas must be in DELTA or DELTA_BLIND!");
    delta = read_col(*cd, 1);
    const Column *ci = find("IVAR");
    if (!ci) throw std::runtime_error("FITS ERROR column IVAR not found");
    ivar = read_col(*ci, 1);
    q.setData(std::move(wave), std::move(delta), std::move(ivar));
    // readAllocResolutionMatrix (io/qso_file.cpp:657-677): N rows of ndiags,
    // transposed to the diagonal storage
    if (const Column *cr = find("RESOMAT")) {
        const long long nd = cr->tdim0 > 0 ? cr->tdim0 : cr->repeat;
        if (nd <= 0 || cr->repeat != nd) throw std::runtime_error("FITS ERROR RESOMAT with inconsistent TDIM");
        const std::vector<double> rows = read_col(*cr, nd);
        auto R = std::make_unique<mxhelp::Resolution>((int)N, (int)nd);
        double *v = R->matrix();
        for (long long i = 0; i < N; ++i)
            for (long long d = 0; d < nd; ++d) v[(size_t)d * N + i] = rows[(size_t)i * nd + d];
        q.pending_rmat = std::move(R);
    }
}

}  // namespace qio
