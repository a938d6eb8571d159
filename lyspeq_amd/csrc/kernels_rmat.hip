// Resolution-matrix mode of the QML engine (gfx950, FP64).
//
// With a resolution matrix R (N x M, banded: row i has rnb values at
// matrix-grid columns rcol0 + i*ros ...) the reference builds, per chunk,
//   Q_k = R Qhat_k R^T,  Qhat_k(v) = (dk/pi) cos(k_c v) sin(dk v/2)/(dk v/2)
//                                    (dk/pi at v = 0),
//   S   = R Shat R^T,    Shat = the 2D signal table at (z, v),
// with v, z taken on the matrix's grid (core/chunk_estimate.cpp:384-480,
// mathtools/matrix_helper.cpp:877-902,1075-1084), then applies the z-bin
// weights on the pixel grid.  Here S and every Q_k (k < KBIN_UPP) are
// materialised once per iteration in the chunk's FP64 matrix arena (qb_off)
// by one launch over the whole batch.
//
// The band window is a difference of edge terms: with the band edges
// k_n, k_n+1 (k_c their mean, dk their difference),
//   Qhat_k(v) = P_n+1(v) - P_n(v),   P_e(v) = sin(k_e v) / (pi v)  (k_e/pi at 0),
// the same function as the reference's product form by the sum-to-product
// identity.  So Q_k = Z_k+1 - Z_k with Z_e = R P_e R^T: one sandwich per band
// edge, and each band is the difference of two consecutive ones, taken in
// registers.  With v = u_b - u_a (u = c ln lambda, relative to a point of the
// block's own range so that |u| stays small), sin(k v) = sin(k u_b) cos(k u_a)
// - cos(k u_b) sin(k u_a): sines and cosines are evaluated once per grid point
// and edge, not once per grid pair and band.
//
// The host lays every R row inside the grid (rcol0 >= 0; a diagonal-storage
// matrix gets nelem/2 extrapolated grid points on each side, whose weights
// are zero), so every tap loop below has the same trip count.
//
// k_rm_sandwich<TJ>: one 16 x TJ output tile (I0.., J0..) per block, tiles
// touching the upper triangle only.  The fine rows of its 16 R rows (the a
// range, 15 ros + rnb) and the fine columns of its TJ R columns (the b range,
// (TJ-1) ros + rnb, in passes of 32) are fixed for the block.  Per edge e
// (and first for S):
//   per pass:  Qh (a x 32, LDS) <- P_e(a, b), or Shat
//              T  (16 x b, LDS) <- R_I Qh                  (rnb taps, FP64 FMA)
//   Z[i][j] = sum_t T[i][j ros + t] R[j][t]                (rnb taps, registers)
//   Q_e-1 = Z_e - Z_e-1 written to the upper triangle and mirrored, as
//   Resolution::sandwich fills it.
#include "qml_common.hpp"
#include "qml_kernels.hpp"

namespace qml {

namespace {
constexpr double RM_PI = 3.14159265358979323846;
constexpr int RM_BW = 32;    // fine columns per pass
constexpr int RM_QP = 33;    // row pitch of Qh (doubles)

}  // namespace

template <int TJ, int EM, int NPM>
__global__ __launch_bounds__(256) void k_rm_sandwich(const ChunkDev *cs, const double *rin, const TableDev *tabs,
                                                     BinsDev bins, RmLayout lay, double *mat) {
    constexpr int NO = TJ / 16;                   // outputs per thread
    // EM: Qh entries per thread (rows ar0 + 8 s, s < EM); 8 EM >= 15 ros + rnb.
    // NPM: passes at most ((TJ - 1) ros + rnb <= 32 NPM), their 1/(pi v) kept in registers
    const ChunkDev c = cs[blockIdx.z];
    if (!c.rm) return;
    const int Np = c.Np, N = c.N, M = c.rM, nb = c.rnb, os = c.ros, kbu = c.kbu;
    const int I0 = blockIdx.y * 16, J0 = blockIdx.x * TJ;
    if (I0 >= Np || J0 >= Np || J0 + TJ - 1 < I0) return;
    const int tid = threadIdx.x, ti = tid >> 4, tj = tid & 15;
    const int i = I0 + ti;
    const size_t NN = (size_t)Np * Np;
    double *Qb = mat + c.qb_off;
    if (I0 >= N || J0 >= N) {                     // all padding
        for (int o = 0; o < NO; ++o) {
            const int j = J0 + tj + 16 * o;
            if (j < i) continue;
            for (int k = 0; k <= kbu; ++k) {
                Qb[k * NN + (size_t)i * Np + j] = 0.0;
                Qb[k * NN + (size_t)j * Np + i] = 0.0;
            }
        }
        return;
    }
    const double *R = rin + c.rin_off;            // N x nb
    const double *grid = R + (size_t)N * nb;      // M wavelengths (units of lambda_Lya)
    const TableDev t = tabs[c.table];
    const int amin = c.rcol0 + I0 * os, ra = 15 * os + nb;
    const int bmin = c.rcol0 + J0 * os, rb = (TJ - 1) * os + nb;
    const int npass = (rb + RM_BW - 1) / RM_BW;
    const int npts = ra + RM_BW * npass;          // the a points, then the b points
    const int ref = min(M - 1, (amin + bmin + rb) / 2);

    extern __shared__ double lds[];
    double *Qh = lds;                             // lay.qr x RM_QP
    double *Tb = Qh + lay.qr * RM_QP;             // 16 x lay.tw
    double *Ri = Tb + 16 * lay.tw;                // 16 x lay.nbm
    double *Rj = Ri + 16 * lay.nbm;               // TJ x lay.nbm
    double *ush = Rj + TJ * lay.nbm;              // lay.npts
    double2 *scs = reinterpret_cast<double2 *>(ush + lay.npts);

    for (int e = tid; e < 16 * nb; e += 256) {
        const int r = e / nb, tt = e - r * nb;
        Ri[r * lay.nbm + tt] = (I0 + r < N) ? R[(size_t)(I0 + r) * nb + tt] : 0.0;
    }
    for (int e = tid; e < TJ * nb; e += 256) {
        const int r = e / nb, tt = e - r * nb;
        Rj[r * lay.nbm + tt] = (J0 + r < N) ? R[(size_t)(J0 + r) * nb + tt] : 0.0;
    }
    const double gref = grid[ref];
    for (int p = tid; p < npts; p += 256) {
        const int g = min(M - 1, p < ra ? amin + p : bmin + (p - ra));
        ush[p] = SPEED_OF_LIGHT * log(grid[g] / gref);
    }
    __syncthreads();

    const int q0 = tid & 15;                      // T step: row ti, columns q0, q0 + 16
    const int bb = tid & 31, ar0 = tid >> 5;      // Qh entries: rows ar0 + 8 s, column bb
    const double *Rti = Ri + ti * lay.nbm;
    const double *Qti = Qh + ti * os * RM_QP;
    double zp[NO];
#pragma unroll
    for (int o = 0; o < NO; ++o) zp[o] = 0.0;

    // 1/(pi v) of this thread's entries in every pass (v = u_b - u_a does not
    // depend on the edge); the entry on the shared grid point (v = 0, at most
    // one per pass and column) gets 0 here and k_e/pi written after the rest
    double inv[NPM][EM];
    int dsel[NPM];
#pragma unroll
    for (int p = 0; p < NPM; ++p) {
        dsel[p] = -1;
        const double ub = (p < npass) ? ush[ra + p * RM_BW + bb] : 0.0;
#pragma unroll
        for (int s = 0; s < EM; ++s) {
            const double d = (ub - ush[ar0 + 8 * s]) * RM_PI;
            const bool live = p < npass && ar0 + 8 * s < ra;
            inv[p][s] = (live && d != 0.0) ? 1.0 / d : 0.0;
            if (live && d == 0.0) dsel[p] = s;
        }
    }

    for (int e = -1; e <= kbu; ++e) {
        double kpi = 0.0;
        if (e >= 0) {
            const double ke = bins.ke[e];
            kpi = ke / RM_PI;
            for (int p = tid; p < npts; p += 256) {
                double sv, cv;
                sincos(ke * ush[p], &sv, &cv);
                scs[p] = make_double2(sv, cv);
            }
            __syncthreads();
        }
        // the a points' (sin, cos) at this edge: the same in every pass
        double2 sa[EM];
        if (e >= 0) {
#pragma unroll
            for (int s = 0; s < EM; ++s) sa[s] = scs[ar0 + 8 * s];
        }
#pragma unroll
        for (int p = 0; p < NPM; ++p) {
            if (p >= npass) break;
            const int pc = p * RM_BW + bb;
            if (e < 0) {
                const int gb = min(M - 1, bmin + pc);
                const double lb = grid[gb];
#pragma unroll 4
                for (int s = 0; s < EM; ++s) {
                    const int a = ar0 + 8 * s, ga = min(M - 1, amin + a);
                    double val = 0.0;
                    if (!bins.turn_off_sfid) {
                        double v, z;
                        vz_pair(grid[ga], lb, ga, gb, v, z);
                        val = signal_interp(t, z, v);
                    }
                    Qh[a * RM_QP + bb] = val;
                }
            } else {
                // rows past the a range fill Qh rows that the T step never reads
                const double2 sb = scs[ra + pc];
#pragma unroll
                for (int s = 0; s < EM; ++s)
                    Qh[(ar0 + 8 * s) * RM_QP + bb] = (sb.x * sa[s].y - sb.y * sa[s].x) * inv[p][s];
                if (dsel[p] >= 0) Qh[(ar0 + 8 * dsel[p]) * RM_QP + bb] = kpi;
            }
            __syncthreads();
            double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
            int tt = 0;
            for (; tt + 3 < nb; tt += 4) {
                const double *q = Qti + tt * RM_QP + q0;
                const double w0 = Rti[tt], w1 = Rti[tt + 1], w2 = Rti[tt + 2], w3 = Rti[tt + 3];
                const double a0 = q[0], a1 = q[16], b0 = q[RM_QP], b1 = q[RM_QP + 16];
                const double c0 = q[2 * RM_QP], c1 = q[2 * RM_QP + 16], d0 = q[3 * RM_QP], d1 = q[3 * RM_QP + 16];
                t0 += w0 * a0;
                t1 += w0 * a1;
                t2 += w1 * b0;
                t3 += w1 * b1;
                t0 += w2 * c0;
                t1 += w2 * c1;
                t2 += w3 * d0;
                t3 += w3 * d1;
            }
            for (; tt + 1 < nb; tt += 2) {
                const double w0 = Rti[tt], w1 = Rti[tt + 1];
                const double *q = Qti + tt * RM_QP + q0;
                t0 += w0 * q[0];
                t1 += w0 * q[16];
                t2 += w1 * q[RM_QP];
                t3 += w1 * q[RM_QP + 16];
            }
            if (tt < nb) {
                const double w0 = Rti[tt];
                const double *q = Qti + tt * RM_QP + q0;
                t0 += w0 * q[0];
                t1 += w0 * q[16];
            }
            Tb[ti * lay.tw + p * RM_BW + q0] = t0 + t2;
            Tb[ti * lay.tw + p * RM_BW + q0 + 16] = t1 + t3;
            __syncthreads();
        }
        double z[NO];
#pragma unroll
        for (int o = 0; o < NO; ++o) {
            const double *tr = Tb + ti * lay.tw + (tj + 16 * o) * os;
            const double *rr = Rj + (tj + 16 * o) * lay.nbm;
            double s0 = 0.0, s1 = 0.0;
            int tt = 0;
            for (; tt + 3 < nb; tt += 4) {
                const double x0 = tr[tt], x1 = tr[tt + 1], x2 = tr[tt + 2], x3 = tr[tt + 3];
                const double y0 = rr[tt], y1 = rr[tt + 1], y2 = rr[tt + 2], y3 = rr[tt + 3];
                s0 += x0 * y0;
                s1 += x1 * y1;
                s0 += x2 * y2;
                s1 += x3 * y3;
            }
            for (; tt + 1 < nb; tt += 2) {
                s0 += tr[tt] * rr[tt];
                s1 += tr[tt + 1] * rr[tt + 1];
            }
            if (tt < nb) s0 += tr[tt] * rr[tt];
            z[o] = s0 + s1;
        }
        if (e != 0) {
            double *Q = Qb + (size_t)(e < 0 ? kbu : e - 1) * NN;
#pragma unroll
            for (int o = 0; o < NO; ++o) {
                const int j = J0 + tj + 16 * o;
                if (j >= i) {
                    const double val = (e < 0) ? z[o] : z[o] - zp[o];
                    Q[(size_t)i * Np + j] = val;
                    Q[(size_t)j * Np + i] = val;
                }
            }
        }
        if (e >= 0) {
#pragma unroll
            for (int o = 0; o < NO; ++o) zp[o] = z[o];
        }
    }
}

RmLayout rmat_layout(int tj, int ra_max, int rb_max, int nb_max) {
    RmLayout l;
    const int npass = (rb_max + RM_BW - 1) / RM_BW;
    l.em = (ra_max <= 80) ? 10 : RM_ROWS / 8;
    l.npass = npass;
    l.qr = 8 * l.em;
    l.tw = RM_BW * npass + 1;
    l.nbm = nb_max;
    // every entry row 8 em - 1 reads a point inside the arrays; even count: (sin, cos) pairs follow
    l.npts = (std::max(ra_max + RM_BW * npass, 8 * l.em) + 1) / 2 * 2;
    l.tj = tj;
    l.bytes = 8 * (l.qr * RM_QP + 16 * l.tw + (16 + tj) * l.nbm + 3 * l.npts);
    return l;
}

bool rmat_layout_ok(const RmLayout &lay) {
    return lay.bytes <= 65536 && ((lay.tj == 32 && lay.em == 10 && lay.npass <= 4) ||
                                  (lay.tj == 16 && (lay.em == 10 || lay.em == RM_ROWS / 8) && lay.npass <= 3));
}

bool launch_rmat(const LaunchCtx &L) {
    if (L.rm_kbu < 0) return true;
    const RmLayout &lay = L.rm_lay;
    const dim3 g(L.np_max / lay.tj, L.np_max / 16, L.nchunks);
    if (lay.tj == 32 && lay.em == 10 && lay.npass <= 4)
        k_rm_sandwich<32, 10, 4><<<g, 256, lay.bytes, L.stream>>>(L.chunks, L.rin, L.tables, L.bins, lay, L.mat);
    else if (lay.tj == 16 && lay.em == 10 && lay.npass <= 3)
        k_rm_sandwich<16, 10, 3><<<g, 256, lay.bytes, L.stream>>>(L.chunks, L.rin, L.tables, L.bins, lay, L.mat);
    else if (lay.tj == 16 && lay.em == RM_ROWS / 8 && lay.npass <= 3)
        k_rm_sandwich<16, RM_ROWS / 8, 3><<<g, 256, lay.bytes, L.stream>>>(L.chunks, L.rin, L.tables, L.bins, lay, L.mat);
    else
        return false;
    return true;
}

}  // namespace qml
