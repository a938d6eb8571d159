// The part of the reference's io/qso_file.hpp (and of mxhelp::Resolution,
// mathtools/matrix_helper.hpp:219-264) that the drop-in Chunk /
// OneQSOEstimate use: the same names and meanings.  In the reference tree
// the reference's own classes are used instead.  This version reads the
// binary spectrum format (io/qso_file.cpp:332-369) and picca's FITS delta
// files (io/qso_file.cpp:376-677); other readers come in through
// qio::QSOFile::setReader.
#pragma once

#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "core/global_numbers.hpp"

namespace mxhelp {
// A spectrum's resolution matrix: diagonal storage (ndiags x size, diagonal
// d holds R[c - (ndiags/2 - d)][c] at column c) or oversampled rows (size x
// nelem, on a grid of ncols = size * osamp + nelem - 1 points).
class Resolution {
    bool is_dia_matrix;
    int nrows, ncols, nelem, osamp;
    std::vector<double> values;

public:
    Resolution(int nm, int ndia);   // diagonal storage, nm x nm with ndia diagonals
    Resolution(const Resolution *rmaster, int i1, int i2);
    int getNCols() const { return ncols; }
    bool isDiaMatrix() const { return is_dia_matrix; }
    double *matrix() const { return const_cast<double *>(values.data()); }
    int getSize() const { return (int)values.size(); }
    int getNElemPerRow() const { return nelem; }
    void cutBoundary(int i1, int i2);
    void deconvolve(double m);
    void oversample(int osamp, double dlambda);
    double getMinMemUsage() const { return process::getMemoryMB((int)values.size()); }
};
}  // namespace mxhelp

namespace qio {

class QSOFile {
    ifileformat PB;
    std::vector<double> wave_, delta_, ivar_;
    int arr_size = 0, _fullsize = 0, shift = 0, num_masked_pixels = 0;
    void _cutMaskedBoundary();
    void _countMaskedPixels();

public:
    // A reader for files that are not in the binary format: fills the
    // parameters and the data of q (readParameters' and readData's fields:
    // arrays, id, z_qso, snr, R_fwhm, dv_kms, dlambda, and optionally Rmat
    // in diagonal storage).
    using Reader = std::function<void(const std::string &fname, QSOFile &q)>;
    static void setReader(Reader r);

    std::string fname;
    double z_qso = 0, snr = 0, dv_kms = -1, R_kms = 0, dlambda = -1, ra = 0, dec = 0;
    long id = 0;
    int R_fwhm = 0, expid = -1, night = -1, fiber = -1, petal = -1;
    std::unique_ptr<mxhelp::Resolution> Rmat;
    // what readAllocResolutionMatrix returns for a Picca file (set by a reader)
    std::unique_ptr<mxhelp::Resolution> pending_rmat;

    QSOFile(const std::string &fname_qso, ifileformat p_or_b);
    // the "copy" constructor also cuts masked boundaries
    QSOFile(const QSOFile &qmaster, int i1, int i2);
    QSOFile(QSOFile &&rhs) = delete;
    QSOFile(const QSOFile &rhs) = delete;

    void closeFile() {}
    int size() const { return arr_size; }
    int realSize() const { return arr_size - num_masked_pixels; }
    double getLengthV() const;
    double *wave() const { return const_cast<double *>(wave_.data()) + shift; }
    double *delta() const { return const_cast<double *>(delta_.data()) + shift; }
    double *ivar() const { return const_cast<double *>(ivar_.data()) + shift; }
    double getMinMemUsage() const;

    void recalcDvDLam();
    void readParameters();
    void readData();
    void cutBoundary(double z_lower_edge, double z_upper_edge);
    int maskOutliers(double factor = 5.0);
    void readMinMaxMedRedshift(double &zmin, double &zmax, double &zmed);
    void readAllocResolutionMatrix();
    // fills the arrays directly (a reader's or a test's data)
    void setData(std::vector<double> wave, std::vector<double> delta, std::vector<double> ivar);
};

// Fills q from HDU n + 1 of a picca delta file named "path[n]" (HDU 2 for a
// bare path; gzip-compressed files too): the header keys, the wavelength /
// delta / ivar columns and, if present, RESOMAT as pending_rmat.
void readPiccaFile(const std::string &fname_qso, QSOFile &q);

}  // namespace qio
