"""k/z binning of the reference (core/global_numbers.cpp:209-471).

Pins: the k edges/centres and z-bin lookup are pinned by the golden detailed
spectrum file (tests/truth/test_it1_quadratic_power_estimate_detailed.dat:
columns z, k1, k2, kc); ``zbin_triangular_first`` by q0_matrix.txt;
``zbin_triangular_first`` and ``zbin_triangular_last`` by the golden Fisher
matrix; the interior triangle and the top hat by their closed forms and by the
partition of unity sum_m w_m(z) = 1 (tests/test_oracle_golden.py::test_redshift_weights_partition_of_unity).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

SPEED_OF_LIGHT = 299792.458
LYA_REST = 1215.67
MY_PI = 3.14159265358979323846
ONE_SIGMA_2_FWHM = 2.35482004503
DOUBLE_EPSILON = 1e-15

TOPHAT, TRIANGULAR = 0, 1


@dataclass
class Bins:
    """bins:: and the specifics:: values the chunk path reads."""
    k0: float
    nklin: int
    dklin: float
    nklog: int
    dklog: float
    z0: float
    z_bin_width: float
    nz: int
    klast: float = -1.0
    method: int = TRIANGULAR
    # specifics (core/global_numbers.cpp:78-168)
    cont_logl_order: int = 1
    cont_lam_order: int = -1
    turn_off_sfid: bool = False
    max_forest_length_v: float = 3.5e4
    number_of_chunks: int = 1
    max_pixel_length_v: float = 150.0
    use_precomputed_fisher: bool = False
    # RedshiftGrowthOn (core/global_numbers.cpp:127) and its exponent
    # FiducialRedshiftPower (core/fiducial_cosmology.cpp:204; compiled-in
    # PD13 value 3.591244 at :268-271)
    redshift_growth_on: bool = False
    fiducial_redshift_power: float = 3.591244
    kband_edges: np.ndarray = field(init=False)
    kband_centers: np.ndarray = field(init=False)
    zbin_centers: np.ndarray = field(init=False)

    def __post_init__(self):
        # setUpBins, core/global_numbers.cpp:209-251
        nk = self.nklin + self.nklog
        ktemp = (self.k0 + self.dklin * self.nklin) * math.pow(10, self.nklog * self.dklog)
        if self.klast > ktemp:
            nk += 1
        edges = [self.k0 + self.dklin * i for i in range(self.nklin + 1)]
        for i in range(1, self.nklog + 1):
            edges.append(edges[self.nklin] * math.pow(10.0, i * self.dklog))
        if self.klast > ktemp:
            edges.append(self.klast)
        self.kband_edges = np.array(edges)
        self.kband_centers = np.array([(edges[i] + edges[i + 1]) / 2.0 for i in range(nk)])
        self.zbin_centers = np.array([self.z0 + self.z_bin_width * zm for zm in range(self.nz)])

    @property
    def nk(self) -> int:
        return len(self.kband_centers)

    @property
    def total_kz_bins(self) -> int:
        return self.nk * self.nz

    @property
    def z_lower_edge(self) -> float:
        return self.zbin_centers[0] - self.z_bin_width / 2.0

    @property
    def z_upper_edge(self) -> float:
        return self.zbin_centers[self.nz - 1] + self.z_bin_width / 2.0

    @property
    def cont_nvecs(self) -> int:
        # calcNvecs, core/global_numbers.cpp:78-90
        if self.cont_logl_order >= 0 or self.cont_lam_order >= 0:
            n = 1
            if self.cont_logl_order > 0:
                n += self.cont_logl_order
            if self.cont_lam_order > 0:
                n += self.cont_lam_order
            return n
        return 0

    @property
    def max_forest_length_per_chunk(self) -> float:
        return self.max_forest_length_v / max(1, self.number_of_chunks)

    @property
    def min_forest_length_v(self) -> float:
        # core/global_numbers.cpp:104-107
        return 0.90 / 2 * self.max_forest_length_per_chunk

    def find_redshift_bin(self, z: float) -> int:
        # core/global_numbers.cpp:347-357
        if z < self.z_lower_edge:
            return -1
        r = int((z - self.z_lower_edge) / self.z_bin_width)
        if r >= self.nz:
            r = self.nz
        return r

    def max_k_index(self, knyq: float) -> int:
        # _getMaxKindex, core/chunk_estimate.cpp:32-39 (std::lower_bound)
        return int(np.searchsorted(self.kband_centers, 0.85 * knyq, side="left"))


def zbin_tophat(z: np.ndarray, zc: float, dz: float) -> np.ndarray:
    """zBinTopHat, core/global_numbers.cpp:439-448: 1 on [zc-dz/2, zc+dz/2]."""
    lo, hi = zc - dz / 2, zc + dz / 2
    return np.where((z >= lo) & (z <= hi), 1.0, 0.0)


def zbin_triangular_first(z: np.ndarray, zc: float, dz: float) -> np.ndarray:
    """zBinTriangular1 (bin 0), core/global_numbers.cpp:393-407."""
    hi = zc + dz
    w = 1.0 - (z - zc) / dz
    return np.where(z < zc, 1.0, np.where(z <= hi, w, 0.0))


def zbin_triangular_last(z: np.ndarray, zc: float, dz: float) -> np.ndarray:
    """zBinTriangular2 (last bin), core/global_numbers.cpp:410-424."""
    lo = zc - dz
    w = 1.0 - (zc - z) / dz
    return np.where(z < lo, 0.0, np.where(z <= zc, w, 1.0))


def zbin_triangular_interior(z: np.ndarray, zc: float, dz: float) -> np.ndarray:
    """zBinTriangular, core/global_numbers.cpp:427-437."""
    lo, hi = zc - dz, zc + dz
    w = 1.0 - np.abs(z - zc) / dz
    return np.where((z >= lo) & (z <= hi), w, 0.0)


def redshift_weight(bins: Bins, z: np.ndarray, zm: int) -> np.ndarray:
    """setRedshiftBinningFunction + the selected function
    (core/global_numbers.cpp:452-471), evaluated elementwise."""
    zc, dz = bins.zbin_centers[zm], bins.z_bin_width
    if bins.method == TOPHAT or bins.nz == 1:
        return zbin_tophat(z, zc, dz)
    if zm == 0:
        return zbin_triangular_first(z, zc, dz)
    if zm == bins.nz - 1:
        return zbin_triangular_last(z, zc, dz)
    return zbin_triangular_interior(z, zc, dz)


def fiducial_pd13(k, z, A, n, alpha, B, beta, lam):
    """Palanque_Delabrouille_etal_2013_fit, core/fiducial_cosmology.cpp:273-282."""
    K_0, Z_0 = 0.009, 3.0
    q0 = np.asarray(k) / K_0 + DOUBLE_EPSILON
    x0 = (1.0 + z) / (1.0 + Z_0)
    lq = np.log(q0)
    return (A * MY_PI / K_0 * np.power(q0, 2.0 + n + alpha * lq + beta * np.log(x0))
            * np.power(x0, B) / (1.0 + lam * np.asarray(k) ** 2))
