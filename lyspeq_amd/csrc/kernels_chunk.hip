// Per-chunk kernels of the QML engine other than the dense MFMA products:
// covariance assembly, Cholesky factorisation, triangular inverse, the FP64
// solves with continuum marginalisation, the d_k pass, generation of the
// projected, noise-weighted Q_k stack, and the reductions.
//
// Notation (per chunk): N pixels padded to Np; D = diag(sqrt(ivar));
// A = I + D (S + sum theta_k Q_k) D = L L^T; Li = L^-1; U = orthonormal
// continuum templates; C^-1 = D Li^T Li D; W = C^-1 U (U^T C^-1 U)^-1;
// K = I - U W^T; C_marg^-1 = K^T C^-1 K.  The stack members are
// Q_k (stored order), H = N' (noise, ivar^-1 on unmasked pixels) and T = S.
// Member j is turned into X_j = D (K Q_j K^T) D in the compute precision;
// the MFMA kernels then form G_j = Li X_j Li^T and the Gram matrix
// E_jj' = <G_j, G_j'> = Tr(C_marg^-1 Q_j C_marg^-1 Q_j').
#include "qml_common.hpp"
#include "qml_kernels.hpp"

namespace qml {

// ---------------------------------------------------------------- assembly --
// A = I + D C D and S, C = S_fid + sum_k theta_k Q_k, on upper 32x32 tiles,
// mirrored (core/chunk_estimate.cpp:384-536,668-698).
__global__ __launch_bounds__(256) void k_assemble(const ChunkDev *cs, const double *in,
                                                  const TableDev *tabs, BinsDev bins,
                                                  const double *theta, double *mat) {
    const ChunkDev c = cs[blockIdx.y];
    const int n32 = c.Np / 32;
    const int I = blockIdx.x / n32, J = blockIdx.x % n32;
    if (blockIdx.x >= n32 * n32 || I > J) return;
    const TableDev t = tabs[c.table];
    const double *wave = in + c.in_off + IN_WAVE * (size_t)c.Np;
    const double *D = in + c.in_off + IN_D * (size_t)c.Np;
    double *A = mat + c.mat_off;
    double *S = A + 3 * (size_t)c.Np * c.Np;
    const int nk = bins.nk;
    for (int e = threadIdx.x; e < 1024; e += 256) {
        const int i = I * 32 + e / 32, j = J * 32 + e % 32;
        if (j < i) continue;
        double a, s;
        if (i >= c.N || j >= c.N) {
            a = (i == j) ? 1.0 : 0.0;
            s = 0.0;
        } else {
            double v, z;
            vz_pair(wave[i], wave[j], i, j, v, z);
            s = bins.turn_off_sfid ? 0.0 : signal_interp(t, z, v);
            const double xx = (v - t.v1) / t.dv;
            const int iv = interp_index(xx, t.nv);
            const double dn = xx - iv;
            const double *r0 = t.q + (size_t)iv * nk, *r1 = r0 + nk;
            double cij = s;
            for (int iz = 0; iz < c.nzq; ++iz) {
                const int zm = c.zm0 + iz;
                const double w = zweight(bins, z, zm);
                if (w == 0.0) continue;
                const double *th = theta + (size_t)zm * nk;
                for (int kn = 0; kn < c.kbu; ++kn) {
                    const double q = r0[kn] * (1 - dn) + r1[kn] * dn;
                    cij += th[kn] * (q * w);
                }
            }
            a = D[i] * cij * D[j] + ((i == j) ? 1.0 : 0.0);
        }
        A[(size_t)i * c.Np + j] = a;
        A[(size_t)j * c.Np + i] = a;
        S[(size_t)i * c.Np + j] = s;
        S[(size_t)j * c.Np + i] = s;
    }
}

// ---------------------------------------------------------------- Cholesky --
// Right-looking blocked Cholesky A = L L^T, one workgroup per chunk, panel
// width 32, trailing update in 64x64 tiles.  L has zeros above the diagonal.
__global__ __launch_bounds__(256) void k_cholesky(const ChunkDev *cs, double *mat, int *status) {
    const ChunkDev c = cs[blockIdx.x];
    const int Np = c.Np, tid = threadIdx.x;
    const double *A = mat + c.mat_off;
    double *L = mat + c.mat_off + (size_t)Np * Np;
    __shared__ double Ld[32][33];
    __shared__ double Pr[64][33];
    __shared__ double Pc[64][33];
    __shared__ double Pan[256][33];
    __shared__ int fail;
    if (tid == 0) fail = 0;
    for (size_t idx = tid; idx < (size_t)Np * Np; idx += 256) {
        const int r = (int)(idx / Np), col = (int)(idx % Np);
        L[idx] = (col <= r) ? A[idx] : 0.0;
    }
    __syncthreads();
    for (int p = 0; p < Np; p += 32) {
        for (int e = tid; e < 1024; e += 256) {
            const int r = e / 32, q = e % 32;
            Ld[r][q] = (q <= r) ? L[(size_t)(p + r) * Np + p + q] : 0.0;
        }
        __syncthreads();
        for (int k = 0; k < 32; ++k) {
            if (tid == 0) {
                const double d = Ld[k][k];
                if (!(d > 0.0)) fail = 1;
                Ld[k][k] = (d > 0.0) ? sqrt(d) : 1.0;
            }
            __syncthreads();
            if (tid > k && tid < 32) Ld[tid][k] /= Ld[k][k];
            __syncthreads();
            for (int e = tid; e < 1024; e += 256) {
                const int r = e / 32, q = e % 32;
                if (r > k && q > k && q <= r) Ld[r][q] -= Ld[r][k] * Ld[q][k];
            }
            __syncthreads();
        }
        for (int e = tid; e < 1024; e += 256) {
            const int r = e / 32, q = e % 32;
            if (q <= r) L[(size_t)(p + r) * Np + p + q] = Ld[r][q];
        }
        // panel: L[r, p:p+32] = A[r, p:p+32] Lpp^-T, one row per thread in LDS
        __syncthreads();
        for (int rb = p + 32; rb < Np; rb += 256) {
            const int r = rb + tid;
            if (r < Np) {
                double *row = L + (size_t)r * Np + p;
                for (int q = 0; q < 32; ++q) Pan[tid][q] = row[q];
#pragma unroll 1
                for (int q = 0; q < 32; ++q) {
                    double s = Pan[tid][q];
#pragma unroll 4
                    for (int k = 0; k < q; ++k) s -= Pan[tid][k] * Ld[q][k];
                    Pan[tid][q] = s / Ld[q][q];
                }
                for (int q = 0; q < 32; ++q) row[q] = Pan[tid][q];
            }
        }
        __syncthreads();
        // trailing update of the lower part of rows/cols >= p + 32
        const int q0 = p + 32;
        const int nrem = Np - q0;
        const int nt = (nrem + 63) / 64;
        for (int ti = 0; ti < nt; ++ti) {
            for (int tj = 0; tj <= ti; ++tj) {
                const int r0 = q0 + ti * 64, c0 = q0 + tj * 64;
                for (int e = tid; e < 64 * 32; e += 256) {
                    const int rr = e / 32, k = e % 32;
                    Pr[rr][k] = (r0 + rr < Np) ? L[(size_t)(r0 + rr) * Np + p + k] : 0.0;
                    Pc[rr][k] = (c0 + rr < Np) ? L[(size_t)(c0 + rr) * Np + p + k] : 0.0;
                }
                __syncthreads();
                const int rr = (tid / 16) * 4, cc = (tid % 16) * 4;
                double acc[4][4] = {};
#pragma unroll 2
                for (int k = 0; k < 32; ++k) {
                    double a[4], b[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) { a[u] = Pr[rr + u][k]; b[u] = Pc[cc + u][k]; }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int w = 0; w < 4; ++w) acc[u][w] += a[u] * b[w];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int r = r0 + rr + u;
#pragma unroll
                    for (int w = 0; w < 4; ++w) {
                        const int col = c0 + cc + w;
                        if (r < Np && col <= r) L[(size_t)r * Np + col] -= acc[u][w];
                    }
                }
                __syncthreads();
            }
        }
    }
    if (tid == 0) status[c.gidx] = fail;
}

// ------------------------------------------------------- triangular inverse --
// Li = L^-1 by 32x32 blocks: Li_ii = inv(L_ii); Li_ij = -Li_ii sum_{k=j}^{i-1} L_ik Li_kj.
__global__ __launch_bounds__(256) void k_trinv(const ChunkDev *cs, double *mat) {
    const ChunkDev c = cs[blockIdx.x];
    const int Np = c.Np, tid = threadIdx.x, nb = Np / 32;
    const double *L = mat + c.mat_off + (size_t)Np * Np;
    double *Li = mat + c.mat_off + 2 * (size_t)Np * Np;
    __shared__ double Lt[32][33];
    __shared__ double Rt[32][33];
    __shared__ double Dt[32][33];
    __shared__ double Tt[32][33];
    for (int i = 0; i < nb; ++i) {
        // invert the diagonal block
        for (int e = tid; e < 1024; e += 256) {
            const int r = e / 32, q = e % 32;
            Lt[r][q] = L[(size_t)(i * 32 + r) * Np + i * 32 + q];
            Dt[r][q] = 0.0;
        }
        __syncthreads();
        if (tid < 32) {
            const int col = tid;
            for (int r = col; r < 32; ++r) {
                double s = (r == col) ? 1.0 : 0.0;
                for (int k = col; k < r; ++k) s -= Lt[r][k] * Dt[k][col];
                Dt[r][col] = s / Lt[r][r];
            }
        }
        __syncthreads();
        for (int e = tid; e < 1024; e += 256) {
            const int r = e / 32, q = e % 32;
            Li[(size_t)(i * 32 + r) * Np + i * 32 + q] = Dt[r][q];
        }
        // zero the blocks right of the diagonal in this block row
        for (size_t e = tid; e < (size_t)32 * (Np - (i + 1) * 32); e += 256) {
            const int r = (int)(e / (Np - (i + 1) * 32)), q = (int)(e % (Np - (i + 1) * 32));
            Li[(size_t)(i * 32 + r) * Np + (i + 1) * 32 + q] = 0.0;
        }
        // off-diagonal blocks of block row i
        for (int j = 0; j < i; ++j) {
            const int a = tid / 8, b0 = (tid % 8) * 4;
            double acc[4] = {0, 0, 0, 0};
            for (int k = j; k < i; ++k) {
                __syncthreads();
                for (int e = tid; e < 1024; e += 256) {
                    const int r = e / 32, q = e % 32;
                    Lt[r][q] = L[(size_t)(i * 32 + r) * Np + k * 32 + q];
                    Rt[r][q] = Li[(size_t)(k * 32 + r) * Np + j * 32 + q];
                }
                __syncthreads();
                for (int q = 0; q < 32; ++q) {
                    const double l = Lt[a][q];
#pragma unroll
                    for (int u = 0; u < 4; ++u) acc[u] += l * Rt[q][b0 + u];
                }
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < 4; ++u) Tt[a][b0 + u] = acc[u];
            __syncthreads();
            double out[4] = {0, 0, 0, 0};
            for (int q = 0; q <= a; ++q) {
                const double d = Dt[a][q];
#pragma unroll
                for (int u = 0; u < 4; ++u) out[u] -= d * Tt[q][b0 + u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) Li[(size_t)(i * 32 + a) * Np + j * 32 + b0 + u] = out[u];
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------- small helpers --
__device__ double block_sum(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    const double s = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    return s;
}

// out = Li v (lower), one wave per row.
__device__ void lower_mv(const double *Li, int Np, const double *v, double *out) {
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    for (int i = w; i < Np; i += 4) {
        const double *row = Li + (size_t)i * Np;
        double s = 0.0;
        for (int k = l; k <= i; k += 64) s += row[k] * v[k];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
        if (l == 0) out[i] = s;
    }
    __syncthreads();
}

// out = Li^T v, one thread per column.
__device__ void lower_tmv(const double *Li, int Np, const double *v, double *out) {
    for (int k = threadIdx.x; k < Np; k += 256) {
        double s = 0.0;
        for (int i = k; i < Np; ++i) s += Li[(size_t)i * Np + k] * v[i];
        out[k] = s;
    }
    __syncthreads();
}

// out = A v (dense), one wave per row.
__device__ void dense_mv(const double *A, int Np, const double *v, double *out) {
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    for (int i = w; i < Np; i += 4) {
        const double *row = A + (size_t)i * Np;
        double s = 0.0;
        for (int k = l; k < Np; k += 64) s += row[k] * v[k];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
        if (l == 0) out[i] = s;
    }
    __syncthreads();
}

// ------------------------------------------------------------------ solves --
// x = C^-1 delta (one FP64 refinement step on A), W = C^-1 U (U^T C^-1 U)^-1,
// y = C_marg^-1 delta = x - W U^T x.
__global__ __launch_bounds__(256) void k_solve(const ChunkDev *cs, const double *in,
                                               const double *mat, double *vec) {
    const ChunkDev c = cs[blockIdx.x];
    const int Np = c.Np, tid = threadIdx.x;
    const double *A = mat + c.mat_off;
    const double *Li = A + 2 * (size_t)Np * Np;
    const double *D = in + c.in_off + IN_D * (size_t)Np;
    const double *delta = in + c.in_off + IN_DELTA * (size_t)Np;
    double *V = vec + c.vec_off;
    double *x = V + V_X * (size_t)Np, *y = V + V_Y * (size_t)Np;
    double *t0 = V + V_T0 * (size_t)Np, *t1 = V + V_T1 * (size_t)Np, *t2 = V + V_T2 * (size_t)Np;
    __shared__ double red[4];
    __shared__ double M[MAX_NU][MAX_NU], Mi[MAX_NU][MAX_NU], ux[MAX_NU];

    // u = A^-1 (D delta), refined once
    for (int i = tid; i < Np; i += 256) t0[i] = D[i] * delta[i];
    __syncthreads();
    lower_mv(Li, Np, t0, t1);
    lower_tmv(Li, Np, t1, x);           // x holds u
    dense_mv(A, Np, x, t2);             // A u
    for (int i = tid; i < Np; i += 256) t2[i] = t0[i] - t2[i];
    __syncthreads();
    lower_mv(Li, Np, t2, t1);
    lower_tmv(Li, Np, t1, t2);
    for (int i = tid; i < Np; i += 256) x[i] = D[i] * (x[i] + t2[i]);
    __syncthreads();

    const int nu = c.nu;
    // C^-1 U_t into the W slots
    for (int t = 0; t < nu; ++t) {
        const double *Ut = in + c.in_off + (IN_U0 + t) * (size_t)Np;
        double *Wt = V + (V_W0 + t) * (size_t)Np;
        for (int i = tid; i < Np; i += 256) t0[i] = D[i] * Ut[i];
        __syncthreads();
        lower_mv(Li, Np, t0, t1);
        lower_tmv(Li, Np, t1, t2);
        for (int i = tid; i < Np; i += 256) Wt[i] = D[i] * t2[i];
        __syncthreads();
    }
    for (int t = 0; t < nu; ++t) {
        const double *Ut = in + c.in_off + (IN_U0 + t) * (size_t)Np;
        for (int s = 0; s < nu; ++s) {
            const double *Ws = V + (V_W0 + s) * (size_t)Np;
            double p = 0.0;
            for (int i = tid; i < Np; i += 256) p += Ut[i] * Ws[i];
            p = block_sum(p, red);
            if (tid == 0) M[t][s] = p;
        }
        double p = 0.0;
        for (int i = tid; i < Np; i += 256) p += Ut[i] * x[i];
        p = block_sum(p, red);
        if (tid == 0) ux[t] = p;
    }
    __syncthreads();
    if (tid == 0 && nu > 0) {
        // Gauss-Jordan on the small SPD matrix
        double a[MAX_NU][2 * MAX_NU];
        for (int r = 0; r < nu; ++r)
            for (int q = 0; q < 2 * nu; ++q) a[r][q] = (q < nu) ? M[r][q] : ((q - nu == r) ? 1.0 : 0.0);
        for (int col = 0; col < nu; ++col) {
            int piv = col;
            for (int r = col + 1; r < nu; ++r)
                if (fabs(a[r][col]) > fabs(a[piv][col])) piv = r;
            for (int q = 0; q < 2 * nu; ++q) { double tmp = a[col][q]; a[col][q] = a[piv][q]; a[piv][q] = tmp; }
            const double d = a[col][col];
            for (int q = 0; q < 2 * nu; ++q) a[col][q] /= d;
            for (int r = 0; r < nu; ++r) {
                if (r == col) continue;
                const double f = a[r][col];
                for (int q = 0; q < 2 * nu; ++q) a[r][q] -= f * a[col][q];
            }
        }
        for (int r = 0; r < nu; ++r)
            for (int q = 0; q < nu; ++q) Mi[r][q] = a[r][nu + q];
    }
    __syncthreads();
    // W = (C^-1 U) Mi, y = x - W (U^T x), elementwise over pixels
    for (int i = tid; i < Np; i += 256) {
        double cu[MAX_NU];
        for (int t = 0; t < nu; ++t) cu[t] = V[(V_W0 + t) * (size_t)Np + i];
        double yi = x[i];
        for (int t = 0; t < nu; ++t) {
            double w = 0.0;
            for (int s = 0; s < nu; ++s) w += cu[s] * Mi[s][t];
            V[(V_W0 + t) * (size_t)Np + i] = w;
            yi -= w * ux[t];
        }
        y[i] = yi;
    }
}

// ------------------------------------------------------------------ d pass --
// P_j[i][0..nu] = sum_b Q_j[i][b] (y_b, W_b) for the stored Q_j and for S,
// Q_j generated in FP64 from the tables.  Rows in blocks of 16, columns in
// tiles of 8 staged through LDS.
constexpr int DP_ROWS = 16, DP_COLS = 8, DP_MEMB = 4;
__global__ __launch_bounds__(256) void k_dpass(const ChunkDev *cs, const double *in,
                                               const TableDev *tabs, BinsDev bins,
                                               const double *vec, double *P) {
    const ChunkDev c = cs[blockIdx.y];
    const int i0 = blockIdx.x * DP_ROWS;
    if (i0 >= c.N) return;
    const int Np = c.Np, tid = threadIdx.x, nk = bins.nk, nu = c.nu;
    const TableDev t = tabs[c.table];
    const double *wave = in + c.in_off + IN_WAVE * (size_t)Np;
    const double *V = vec + c.vec_off;
    __shared__ double sq[DP_ROWS * DP_COLS][MAX_NK];
    __shared__ double sw[8][DP_ROWS][DP_COLS];
    __shared__ double sS[DP_ROWS][DP_COLS];
    __shared__ double sR[DP_COLS][1 + MAX_NU];
    __shared__ int sv[DP_ROWS][DP_COLS];
    __shared__ double sf[DP_ROWS][DP_COLS];
    const int r = tid % DP_ROWS, g = tid / DP_ROWS;   // 16 groups
    const int nmem = c.n + 1;                          // stored Q and S
    const int npass = (nmem + 16 * DP_MEMB - 1) / (16 * DP_MEMB);
    const int ncol = 1 + nu;
    for (int pass = 0; pass < npass; ++pass) {
        double acc[DP_MEMB][1 + MAX_NU];
        for (int u = 0; u < DP_MEMB; ++u)
            for (int q = 0; q <= MAX_NU; ++q) acc[u][q] = 0.0;
        for (int b0 = 0; b0 < c.N; b0 += DP_COLS) {
            __syncthreads();
            if (tid < DP_ROWS * DP_COLS) {
                const int rr = tid / DP_COLS, bb = tid % DP_COLS;
                const int i = i0 + rr, b = b0 + bb;
                int iv = 0;
                double fr = 0.0, s = 0.0;
                const bool ok = (i < c.N && b < c.N);
                double z = 0.0;
                if (ok) {
                    double v;
                    vz_pair(wave[i], wave[b], i, b, v, z);
                    const double xx = (v - t.v1) / t.dv;
                    iv = interp_index(xx, t.nv);
                    fr = xx - iv;
                    s = bins.turn_off_sfid ? 0.0 : signal_interp(t, z, v);
                }
                sv[rr][bb] = iv;
                sf[rr][bb] = fr;
                sS[rr][bb] = s;
                for (int iz = 0; iz < c.nzq; ++iz) sw[iz][rr][bb] = ok ? zweight(bins, z, c.zm0 + iz) : 0.0;
            }
            if (tid < DP_COLS * ncol) {
                const int bb = tid / ncol, q = tid % ncol;
                const int b = b0 + bb;
                double val = 0.0;
                if (b < c.N) val = (q == 0) ? V[V_Y * (size_t)Np + b] : V[(V_W0 + q - 1) * (size_t)Np + b];
                sR[bb][q] = val;
            }
            __syncthreads();
            {
                const int e = tid / 2, h = tid % 2;
                const int rr = e / DP_COLS, bb = e % DP_COLS;
                const int iv = sv[rr][bb];
                const double fr = sf[rr][bb];
                const double *r0 = t.q + (size_t)iv * nk, *r1 = r0 + nk;
                for (int kn = h; kn < c.kbu; kn += 2) sq[e][kn] = r0[kn] * (1 - fr) + r1[kn] * fr;
            }
            __syncthreads();
            for (int bb = 0; bb < DP_COLS; ++bb) {
                double R[1 + MAX_NU];
                for (int q = 0; q < ncol; ++q) R[q] = sR[bb][q];
#pragma unroll
                for (int u = 0; u < DP_MEMB; ++u) {
                    const int j = g + 16 * (pass * DP_MEMB + u);
                    if (j >= nmem) continue;
                    double qv;
                    if (j < c.n) {
                        const int iz = j / c.kbu, kn = j % c.kbu;
                        qv = sq[r * DP_COLS + bb][kn] * sw[iz][r][bb];
                    } else {
                        qv = sS[r][bb];
                    }
                    for (int q = 0; q < ncol; ++q) acc[u][q] += qv * R[q];
                }
            }
        }
        const int i = i0 + r;
        if (i < c.N) {
            for (int u = 0; u < DP_MEMB; ++u) {
                int j = g + 16 * (pass * DP_MEMB + u);
                if (j >= nmem) continue;
                const int slot = (j < c.n) ? j : c.n + 1;      // S is member T
                double *dst = P + c.pj_off + ((size_t)slot * Np + i) * (1 + MAX_NU);
                for (int q = 0; q < ncol; ++q) dst[q] = acc[u][q];
            }
        }
    }
}

// d_j = y^T Q_j y, a_H = N' W, s_j = W^T a_j for all members.
__global__ __launch_bounds__(256) void k_small(const ChunkDev *cs, const double *in,
                                               const double *vec, double *P, double *sj) {
    const ChunkDev c = cs[blockIdx.x];
    const int Np = c.Np, tid = threadIdx.x, nu = c.nu;
    const double *V = vec + c.vec_off;
    const double *y = V + V_Y * (size_t)Np;
    const double *Npr = in + c.in_off + IN_NPRIME * (size_t)Np;
    __shared__ double red[4];
    double *s = sj + c.sj_off;
    double *d = s + (size_t)c.m * MAX_NU * MAX_NU;
    // a_H
    for (int i = tid; i < Np; i += 256) {
        double *dst = P + c.pj_off + ((size_t)c.n * Np + i) * (1 + MAX_NU);
        dst[0] = 0.0;
        for (int q = 0; q < nu; ++q) dst[1 + q] = (i < c.N) ? Npr[i] * V[(V_W0 + q) * (size_t)Np + i] : 0.0;
    }
    // rows beyond N of the other members are never written by k_dpass
    for (int j = 0; j < c.m; ++j) {
        if (j == c.n) continue;
        for (int i = c.N + tid; i < Np; i += 256) {
            double *dst = P + c.pj_off + ((size_t)j * Np + i) * (1 + MAX_NU);
            for (int q = 0; q <= MAX_NU; ++q) dst[q] = 0.0;
        }
    }
    __syncthreads();
    for (int j = 0; j < c.m; ++j) {
        const double *Pj = P + c.pj_off + (size_t)j * Np * (1 + MAX_NU);
        if (j < c.n) {
            double p = 0.0;
            for (int i = tid; i < c.N; i += 256) p += y[i] * Pj[(size_t)i * (1 + MAX_NU)];
            p = block_sum(p, red);
            if (tid == 0) d[j] = p;
        }
        for (int t = 0; t < nu; ++t) {
            const double *Wt = V + (V_W0 + t) * (size_t)Np;
            for (int q = 0; q < nu; ++q) {
                double p = 0.0;
                for (int i = tid; i < c.N; i += 256) p += Wt[i] * Pj[(size_t)i * (1 + MAX_NU) + 1 + q];
                p = block_sum(p, red);
                if (tid == 0) s[(size_t)j * MAX_NU * MAX_NU + t * MAX_NU + q] = p;
            }
        }
    }
}

// ------------------------------------------------------- stack generation --
// X_j = D (K M_j K^T) D for every member, stored in the compute type,
// 32x32 tiles over the full matrix.
template <class OutT>
__global__ __launch_bounds__(256) void k_xgen(const ChunkDev *cs, const double *in,
                                              const TableDev *tabs, BinsDev bins,
                                              const double *P, const double *sj, OutT *X) {
    const ChunkDev c = cs[blockIdx.y];
    const int n32 = c.Np / 32;
    if (blockIdx.x >= n32 * n32) return;
    const int I = blockIdx.x / n32, J = blockIdx.x % n32;
    const int Np = c.Np, nk = bins.nk, nu = c.nu;
    const TableDev t = tabs[c.table];
    const double *wave = in + c.in_off + IN_WAVE * (size_t)Np;
    const double *D = in + c.in_off + IN_D * (size_t)Np;
    const double *Npr = in + c.in_off + IN_NPRIME * (size_t)Np;
    const double *s = sj + c.sj_off;
    OutT *Xc = X + c.stk_off;
    const size_t msz = (size_t)Np * Np;
    for (int e = threadIdx.x; e < 1024; e += 256) {
        const int i = I * 32 + e / 32, b = J * 32 + e % 32;
        const size_t off = (size_t)i * Np + b;
        if (i >= c.N || b >= c.N) {
            for (int j = 0; j < c.m; ++j) Xc[j * msz + off] = (OutT)0;
            continue;
        }
        double v, z;
        vz_pair(wave[i], wave[b], i, b, v, z);
        const double sc = D[i] * D[b];
        const double xx = (v - t.v1) / t.dv;
        const int iv = interp_index(xx, t.nv);
        const double dn = xx - iv;
        const double *r0 = t.q + (size_t)iv * nk, *r1 = r0 + nk;
        double Ui[MAX_NU], Ub[MAX_NU];
        for (int q = 0; q < nu; ++q) {
            Ui[q] = in[c.in_off + (IN_U0 + q) * (size_t)Np + i];
            Ub[q] = in[c.in_off + (IN_U0 + q) * (size_t)Np + b];
        }
        double w[8];
        for (int iz = 0; iz < c.nzq; ++iz) w[iz] = zweight(bins, z, c.zm0 + iz);
        auto corr = [&](int j) {
            if (nu == 0) return 0.0;
            const double *ai = P + c.pj_off + ((size_t)j * Np + i) * (1 + MAX_NU) + 1;
            const double *ab = P + c.pj_off + ((size_t)j * Np + b) * (1 + MAX_NU) + 1;
            const double *sm = s + (size_t)j * MAX_NU * MAX_NU;
            double r = 0.0;
            for (int q = 0; q < nu; ++q) {
                r -= Ui[q] * ab[q] + ai[q] * Ub[q];
                double u = 0.0;
                for (int p = 0; p < nu; ++p) u += sm[q * MAX_NU + p] * Ub[p];
                r += Ui[q] * u;
            }
            return r;
        };
        for (int kn = 0; kn < c.kbu; ++kn) {
            const double q = r0[kn] * (1 - dn) + r1[kn] * dn;
            for (int iz = 0; iz < c.nzq; ++iz) {
                const int j = iz * c.kbu + kn;
                Xc[j * msz + off] = (OutT)(sc * (q * w[iz] + corr(j)));
            }
        }
        const double hval = (i == b) ? Npr[i] : 0.0;
        Xc[c.n * msz + off] = (OutT)(sc * (hval + corr(c.n)));
        const double sval = bins.turn_off_sfid ? 0.0 : signal_interp(t, z, v);
        Xc[(c.n + 1) * msz + off] = (OutT)(sc * (sval + corr(c.n + 1)));
    }
}

// Li in the compute type.
template <class OutT>
__global__ void k_convert_li(const ChunkDev *cs, const double *mat, OutT *lic) {
    const ChunkDev c = cs[blockIdx.y];
    const size_t n = (size_t)c.Np * c.Np;
    const double *Li = mat + c.mat_off + 2 * n;
    for (size_t e = blockIdx.x * 256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256)
        lic[c.lic_off + e] = (OutT)Li[e];
}

// ------------------------------------------------------------ Gram reduce --
// E = sum over splits; per-chunk F, d, b, t in the reference's index space.
__global__ __launch_bounds__(256) void k_gram_finalize(const ChunkDev *cs, const ChunkOut *outs,
                                                       const double *gram, const double *sj,
                                                       const int *status, int nk,
                                                       double *fc, double *dbt) {
    const ChunkDev c = cs[blockIdx.x];
    const ChunkOut o = outs[c.gidx];
    const int tid = threadIdx.x;
    const int ntile = c.mt * (c.mt + 1) / 2;
    const double *d = sj + c.sj_off + (size_t)c.m * MAX_NU * MAX_NU;
    double *F = fc + o.fc_off;
    double *B = dbt + o.dbt_off;
    const int fail = status[c.gidx];
    for (int e = tid; e < o.nq * o.nq; e += 256) F[e] = 0.0;
    for (int e = tid; e < 3 * o.nq; e += 256) B[e] = 0.0;
    __syncthreads();
    if (fail) return;
    auto E = [&](int a, int b) {
        if (a > b) { int tmp = a; a = b; b = tmp; }
        const int I = a / TILE, J = b / TILE;
        const int tidx = I * c.mt - I * (I - 1) / 2 + (J - I);
        const int r = a % TILE, q = b % TILE;
        double sum = 0.0;
        for (int sp = 0; sp < c.ntu; ++sp)
            sum += gram[c.gram_off + ((size_t)sp * ntile + tidx) * TILE * TILE + r * TILE + q];
        return sum;
    };
    auto kz = [&](int sidx) { return (sidx / c.kbu) * nk + (sidx % c.kbu); };
    for (int e = tid; e < c.n * c.n; e += 256) {
        const int a = e / c.n, b = e % c.n;
        if (b < a) continue;
        F[(size_t)kz(a) * o.nq + kz(b)] = E(a, b);
    }
    for (int a = tid; a < c.n; a += 256) {
        B[kz(a)] = d[a];
        B[o.nq + kz(a)] = E(a, c.n);
        B[2 * o.nq + kz(a)] = E(a, c.n + 1);
    }
}

// --------------------------------------------------------------- accumulate --
// fisher_sum[(T+1) fis + j + i T] += F_chunk[j + i nq] for j >= i, and the
// d/b/t vectors, in chunk-id order (core/chunk_estimate.cpp:893-903).
__global__ __launch_bounds__(256) void k_accumulate(const ChunkOut *outs, const int *status, int nchunks,
                                                    int T, const double *fc, const double *dbt,
                                                    double *Fsum, double *dbt_sum) {
    // each thread owns fixed columns of row a, so the sums run in chunk order
    const int a = blockIdx.x;
    for (int b = a + threadIdx.x; b < T; b += 256) {
        double s = Fsum[(size_t)a * T + b];
        for (int ci = 0; ci < nchunks; ++ci) {
            const ChunkOut o = outs[ci];
            if (status[ci] || a < o.fis || b >= o.fis + o.nq) continue;
            s += fc[o.fc_off + (size_t)(a - o.fis) * o.nq + (b - o.fis)];
        }
        Fsum[(size_t)a * T + b] = s;
    }
    if (threadIdx.x < 3) {
        const int k = threadIdx.x;
        double s = dbt_sum[k * T + a];
        for (int ci = 0; ci < nchunks; ++ci) {
            const ChunkOut o = outs[ci];
            if (status[ci] || a < o.fis || a >= o.fis + o.nq) continue;
            s += dbt[o.dbt_off + k * o.nq + (a - o.fis)];
        }
        dbt_sum[k * T + a] = s;
    }
}

// ------------------------------------------------------------------ launch --
void launch_assemble(const LaunchCtx &L, const double *theta) {
    dim3 grid((L.np_max / 32) * (L.np_max / 32), L.nchunks);
    k_assemble<<<grid, 256, 0, L.stream>>>(L.chunks, L.in, L.tables, L.bins, theta, L.mat);
}
void launch_cholesky(const LaunchCtx &L, int *status) {
    k_cholesky<<<L.nchunks, 256, 0, L.stream>>>(L.chunks, L.mat, status);
}
void launch_trinv(const LaunchCtx &L) { k_trinv<<<L.nchunks, 256, 0, L.stream>>>(L.chunks, L.mat); }
void launch_solve(const LaunchCtx &L) { k_solve<<<L.nchunks, 256, 0, L.stream>>>(L.chunks, L.in, L.mat, L.vec); }
void launch_dpass(const LaunchCtx &L) {
    dim3 grid((L.np_max + DP_ROWS - 1) / DP_ROWS, L.nchunks);
    k_dpass<<<grid, 256, 0, L.stream>>>(L.chunks, L.in, L.tables, L.bins, L.vec, L.pj);
}
void launch_small(const LaunchCtx &L) { k_small<<<L.nchunks, 256, 0, L.stream>>>(L.chunks, L.in, L.vec, L.pj, L.sj); }
void launch_xgen(const LaunchCtx &L, void *X, int prec) {
    dim3 grid((L.np_max / 32) * (L.np_max / 32), L.nchunks);
    if (prec == 0)
        k_xgen<double><<<grid, 256, 0, L.stream>>>(L.chunks, L.in, L.tables, L.bins, L.pj, L.sj, (double *)X);
    else
        k_xgen<float><<<grid, 256, 0, L.stream>>>(L.chunks, L.in, L.tables, L.bins, L.pj, L.sj, (float *)X);
}
void launch_convert_li(const LaunchCtx &L, void *lic) {
    dim3 grid(64, L.nchunks);
    k_convert_li<float><<<grid, 256, 0, L.stream>>>(L.chunks, L.mat, (float *)lic);
}
void launch_gram_finalize(const LaunchCtx &L, const ChunkOut *outs, const double *gram, const int *status,
                          double *fc, double *dbt) {
    k_gram_finalize<<<L.nchunks, 256, 0, L.stream>>>(L.chunks, outs, gram, L.sj, status, L.bins.nk, fc, dbt);
}
void launch_accumulate(hipStream_t st, const ChunkOut *outs, const int *status, int nchunks, int T,
                       const double *fc, const double *dbt, double *Fsum, double *dbt_sum) {
    k_accumulate<<<T, 256, 0, st>>>(outs, status, nchunks, T, fc, dbt, Fsum, dbt_sum);
}

}  // namespace qml
