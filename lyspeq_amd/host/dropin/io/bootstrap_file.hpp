// The bootstrap outputs of the reference's io/bootstrap_file.hpp that the
// per-chunk path feeds, with the same names, files and layouts:
//
// * BootstrapChunksFile (io/bootstrap_file.cpp:137-185): <base>-bootchunks-
//   <pe>.fits, one image HDU per chunk written by _saveChunkResults
//   (core/quadratic_estimate.cpp:33-55): d, b, t (ndim each) and then the
//   upper diagonals of the chunk's Fisher matrix, d = 0 .. ndim - 1
//   (mxhelp::getDiagonal), 3 ndim + ndim (ndim + 1) / 2 values; header keys
//   TARGETID, ZQSO, RA, DEC, NQDIM, ISTART.
// * saveBootstrapRealizations / readBootstrapRealizations
//   (io/bootstrap_file.cpp:17-135): <base>-bootstrap-realizations.fits, an
//   image REALIZATIONS of nboots rows of nk * nz powers (keys NBOOTS, NK, NZ,
//   FASTBOOT) and an image SOLV_INVF of the solver's inverse Fisher matrix.
//   The returned name carries cfitsio's "!" (overwrite) prefix, as the
//   reference's does.
//
// The files are plain FITS written without cfitsio: 2880-byte blocks,
// big-endian IEEE doubles, values in the headers printed to round-trip.
#pragma once

#include <cstdio>
#include <memory>
#include <string>

namespace ioh {

std::string saveBootstrapRealizations(const std::string &base, const double *allpowers, const double *invfisher,
                                      unsigned int nboots, int nk, int nz, bool fastbootstrap,
                                      const char *comment);

void readBootstrapRealizations(const std::string &fname, std::unique_ptr<double[]> &allpowers,
                               std::unique_ptr<double[]> &invfisher, unsigned int &nboots, int &nk, int &nz,
                               bool &fastbootstrap);

class BootstrapChunksFile {
    std::FILE *file = nullptr;
    bool any = false;

public:
    BootstrapChunksFile(const std::string &base, int thispe);
    ~BootstrapChunksFile();
    BootstrapChunksFile(const BootstrapChunksFile &) = delete;
    BootstrapChunksFile &operator=(const BootstrapChunksFile &) = delete;

    void writeChunk(const double *pk, const double *nk, const double *tk, const double *fisher, int ndim,
                    int fisher_index_start, long id, double z_qso, double ra, double dec);
};

}  // namespace ioh
