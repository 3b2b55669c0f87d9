// Device code of libpizza_amd, part 3: the collocation method
// (cheb_method = 'colloc', matrix_solve = 'direct').  Fields and time
// arrays hold grid values; the per-m systems are dense and are applied as
// explicit inverses formed on the device whenever the matrices change.
// Included once by pz_api.hip, after pz_mphase.h.
#pragma once
#include "pz_mphase.h"

__device__ __forceinline__ double msub(double a, double b) { return a - b; }
__device__ __forceinline__ cd msub(cd a, cd b) { return csub(a, b); }
// modulus (Fortran abs of a real or complex element)
__device__ __forceinline__ double mabs2(double a) { return fabs(a); }
__device__ __forceinline__ double mabs2(cd a) { return hypot(a.x, a.y); }

// ---------------------------------------------------------------------------
// In-place Gauss-Jordan inversion with partial pivoting of the left half of
// the row-major n x 2n block W = [A | I] by one workgroup; on return the
// right half holds A^{-1}.  f: LDS scratch of n values.  Returns the
// LAPACK-style info (k + 1 of the first zero pivot, 0 if none).
// ---------------------------------------------------------------------------
template <class M>
__device__ int gj_invert(M *W, int n, M *f)
{
    __shared__ int s_p, s_info;
    __shared__ double s_v[32];
    __shared__ int s_i[32];
    const int tid = threadIdx.x, nt = blockDim.x, w2 = 2 * n;
    if (tid == 0) s_info = 0;
    for (int t = tid; t < n * n; t += nt) {
        const int i = t / n, j = t - i * n;
        W[(size_t)i * w2 + n + j] = i == j ? mk_elem<M>(1.0, 0.0) : mzero<M>();
    }
    __syncthreads();
    for (int k = 0; k < n; k++) {
        // pivot: first row of the largest |re| + |im| in column k (i >= k)
        double bv = -1.0;
        int bi = n;
        for (int i = k + tid; i < n; i += nt) {
            const double v = mabs1(W[(size_t)i * w2 + k]);
            if (v > bv) { bv = v; bi = i; }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const double ov = __shfl_xor(bv, o, 64);
            const int oi = __shfl_xor(bi, o, 64);
            if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
        }
        if ((tid & 63) == 0) { s_v[tid >> 6] = bv; s_i[tid >> 6] = bi; }
        __syncthreads();
        if (tid == 0) {
            double v = s_v[0];
            int ii = s_i[0];
            for (int q = 1; q < (nt + 63) / 64; q++)
                if (s_v[q] > v || (s_v[q] == v && s_i[q] < ii)) { v = s_v[q]; ii = s_i[q]; }
            s_p = ii;
            if (!(v > 0.0) && s_info == 0) s_info = k + 1;
        }
        __syncthreads();
        const int p = s_p;
        if (s_info != 0) break;
        if (p != k)
            for (int j = k + tid; j < w2; j += nt) {
                const M t = W[(size_t)k * w2 + j];
                W[(size_t)k * w2 + j] = W[(size_t)p * w2 + j];
                W[(size_t)p * w2 + j] = t;
            }
        __syncthreads();
        const M rp = mrecip(W[(size_t)k * w2 + k]);
        __syncthreads();
        for (int j = k + tid; j < w2; j += nt) W[(size_t)k * w2 + j] = mmm(W[(size_t)k * w2 + j], rp);
        for (int i = tid; i < n; i += nt) f[i] = i == k ? mzero<M>() : W[(size_t)i * w2 + k];
        __syncthreads();
        const int wj = w2 - k;
        for (int t = tid; t < n * wj; t += nt) {
            const int i = t / wj, j = k + (t - i * wj);
            if (i != k) W[(size_t)i * w2 + j] = msub(W[(size_t)i * w2 + j], mmm(f[i], W[(size_t)k * w2 + j]));
        }
        __syncthreads();
    }
    __syncthreads();
    return s_info;
}

#define GJ_BMAX 16 // widest panel (pivot columns per panel)
#define GJ_CH 64   // columns per chunk of the panel update
// dynamic LDS bytes of gj_invert_blk for n x n matrices of element M and
// panels of B pivot columns
template <class M>
__host__ __device__ inline size_t gj_lds_bytes(int n, int B)
{
    return sizeof(M) * (2 * (size_t)n * B + (size_t)B * GJ_CH + 2 * (size_t)B);
}

// ---------------------------------------------------------------------------
// Blocked Gauss-Jordan inversion with partial pivoting of the left half of
// the row-major n x 2n block W = [A | I] by one workgroup, forming only the
// first nr columns of A^{-1} in W's right half.  Each panel of B pivot
// columns is eliminated in LDS with the pivots and multipliers of the
// unblocked elimination (gj_invert).  The panel's row swaps, then its
// composed elimination T = I + U e_K^T, go to the rest of W in one pass:
// U = C (I - Lc)^{-1}, where column q of C holds the multipliers of pivot
// q (c_i = -f_i / p for i /= k, 1/p - 1 at the pivot row) and Lc is the
// strictly lower part of C's pivot rows.  W is read and written n / B
// times instead of n times; B <= GJ_BMAX.  lds: gj_lds_bytes<M>(n, B)
// bytes.  Returns the LAPACK-style info (k + 1 of the first zero pivot, 0
// if none).
// ---------------------------------------------------------------------------
template <class M>
__device__ int gj_invert_blk(M *W, int n, int nr, M *lds, int B)
{
    __shared__ int s_p, s_info, s_piv[GJ_BMAX];
    __shared__ double s_v[32];
    __shared__ int s_i[32];
    M *P = lds, *C = P + (size_t)n * B, *Wk = C + (size_t)n * B, *prow = Wk + B * GJ_CH,
      *coef = prow + B;
    const int tid = threadIdx.x, nt = blockDim.x, w2 = 2 * n;
    if (tid == 0) s_info = 0;
    for (int t = tid; t < n * nr; t += nt) {
        const int i = t / nr, j = t - i * nr;
        W[(size_t)i * w2 + n + j] = i == j ? mk_elem<M>(1.0, 0.0) : mzero<M>();
    }
    __syncthreads();
    for (int k0 = 0; k0 < n; k0 += B) {
        const int b = min(B, n - k0);
        for (int t = tid; t < n * B; t += nt) {
            const int i = t / B, c = t - i * B;
            P[t] = c < b ? W[(size_t)i * w2 + k0 + c] : mzero<M>();
            C[t] = mzero<M>();
        }
        __syncthreads();
        // ---- the panel, pivot by pivot, in LDS
        for (int kk = 0; kk < b; kk++) {
            const int k = k0 + kk;
            double bv = -1.0;
            int bi = n;
            for (int i = k + tid; i < n; i += nt) {
                const double v = mabs1(P[(size_t)i * B + kk]);
                if (v > bv) { bv = v; bi = i; }
            }
            for (int o = 32; o > 0; o >>= 1) {
                const double ov = __shfl_xor(bv, o, 64);
                const int oi = __shfl_xor(bi, o, 64);
                if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
            }
            if ((tid & 63) == 0) { s_v[tid >> 6] = bv; s_i[tid >> 6] = bi; }
            __syncthreads();
            if (tid == 0) {
                double v = s_v[0];
                int ii = s_i[0];
                for (int q = 1; q < (nt + 63) / 64; q++)
                    if (s_v[q] > v || (s_v[q] == v && s_i[q] < ii)) { v = s_v[q]; ii = s_i[q]; }
                s_p = ii;
                s_piv[kk] = ii;
                if (!(v > 0.0) && s_info == 0) s_info = k + 1;
            }
            __syncthreads();
            if (s_info != 0) break;
            const int p = s_p;
            if (p != k && tid < B) {
                const M a = P[(size_t)k * B + tid], c = C[(size_t)k * B + tid];
                P[(size_t)k * B + tid] = P[(size_t)p * B + tid];
                C[(size_t)k * B + tid] = C[(size_t)p * B + tid];
                P[(size_t)p * B + tid] = a;
                C[(size_t)p * B + tid] = c;
            }
            __syncthreads();
            const M r = mrecip(P[(size_t)k * B + kk]);
            if (tid < B) prow[tid] = P[(size_t)k * B + tid];
            for (int i = tid; i < n; i += nt)
                C[(size_t)i * B + kk] =
                    i == k ? msub(r, mk_elem<M>(1.0, 0.0)) : mneg(mmm(P[(size_t)i * B + kk], r));
            __syncthreads();
            for (int t = tid; t < n * B; t += nt) {
                const int i = t / B, c = t - i * B;
                if (c >= kk && c < b) P[t] = madd(P[t], mmm(C[(size_t)i * B + kk], prow[c]));
            }
            __syncthreads();
        }
        if (s_info != 0) break;
        // ---- the other columns J = [k0 + b, n) and [n, n + nr): swaps in order
        const int nl = n - k0 - b, nJ = nl + nr;
        for (int kk = 0; kk < b; kk++) {
            const int k = k0 + kk, p = s_piv[kk];
            if (p != k)
                for (int jj = tid; jj < nJ; jj += nt) {
                    const int j = jj < nl ? k0 + b + jj : n + (jj - nl);
                    const M a = W[(size_t)k * w2 + j];
                    W[(size_t)k * w2 + j] = W[(size_t)p * w2 + j];
                    W[(size_t)p * w2 + j] = a;
                }
            __syncthreads();
        }
        // U = C (I - Lc)^{-1}, in place, last column first
        for (int kk = b - 1; kk >= 0; kk--) {
            if (tid < B) coef[tid] = (tid > kk && tid < b) ? C[(size_t)(k0 + tid) * B + kk] : mzero<M>();
            __syncthreads();
            for (int i = tid; i < n; i += nt) {
                M u = C[(size_t)i * B + kk];
                for (int q = kk + 1; q < b; q++) u = madd(u, mmm(C[(size_t)i * B + q], coef[q]));
                C[(size_t)i * B + kk] = u;
            }
            __syncthreads();
        }
        // W[:, J] += U W[K, J], GJ_CH columns at a time
        for (int j0 = 0; j0 < nJ; j0 += GJ_CH) {
            const int cw = min(GJ_CH, nJ - j0);
            for (int t = tid; t < b * cw; t += nt) {
                const int q = t / cw, jc = t - q * cw, jj = j0 + jc;
                const int j = jj < nl ? k0 + b + jj : n + (jj - nl);
                Wk[q * GJ_CH + jc] = W[(size_t)(k0 + q) * w2 + j];
            }
            __syncthreads();
            for (int t = tid; t < n * cw; t += nt) {
                const int i = t / cw, jc = t - i * cw, jj = j0 + jc;
                const int j = jj < nl ? k0 + b + jj : n + (jj - nl);
                M v = W[(size_t)i * w2 + j];
                for (int q = 0; q < b; q++) v = madd(v, mmm(C[(size_t)i * B + q], Wk[q * GJ_CH + jc]));
                W[(size_t)i * w2 + j] = v;
            }
            __syncthreads();
        }
    }
    __syncthreads();
    return s_info;
}

#define PZ_CM(k, n) p.cM[(size_t)(k) * N + (n)]
#define PZ_CDM(k, n) p.cdM[(size_t)(k) * N + (n)]
#define PZ_CD2M(k, n) p.cd2M[(size_t)(k) * N + (n)]

// Assembly of the collocation matrices of one (field, m) into the left half
// of W (row-major, 2n wide), then the scalings of the reference.
//   T / xi: get_tempMat / get_xiMat (update_temp_coll.f90:310-393), rows
//           preconditioned, fac[r] returned;
//   psi, m /= 0: get_psiMat (update_psi_coll_smat.f90:674-837), rows then
//           columns max-scaled, fac[0..2N) rows and fac[2N..4N) columns;
//   u_phi, m = 0: get_uphiMat (update_psi_coll_smat.f90:919-987).
template <class M>
__device__ void coll_assemble(const PzDev &p, const RadView &R, int fld, int m, double hdif, M *W, double *fac)
{
    const int N = p.N, nc = p.ncheb, tid = threadIdx.x, nt = blockDim.x;
    const double wimp = p.ss->wimp_lin[0], rn = p.rnorm, dm = (double)m, dm2 = dm * dm;
    const bool psi = fld == 2 && m != 0;
    const int n = psi ? 2 * N : N, w2 = 2 * n;
    for (int t = tid; t < n * n; t += nt) {
        const int r = t / n, c = t - r * n;
        M v = mzero<M>();
        if (fld < 2) {
            const double df = (fld == 0 ? p.TdiffFac : p.XidiffFac) * hdif;
            const int ktop = fld == 0 ? p.ktopt : p.ktopxi, kbot = fld == 0 ? p.kbott : p.kbotxi;
            double a;
            if (r == 0) a = c < nc ? rn * (ktop == 1 ? PZ_CM(0, c) : PZ_CDM(0, c)) : 0.0;
            else if (r == N - 1) a = c < nc ? rn * (kbot == 1 ? PZ_CM(N - 1, c) : PZ_CDM(N - 1, c)) : 0.0;
            else
                a = rn * (PZ_CM(r, c) -
                          wimp * df * (PZ_CD2M(r, c) + R.or1(r) * PZ_CDM(r, c) - dm2 * R.or2(r) * PZ_CM(r, c)));
            if (c == 0 || c == N - 1) a *= p.boundary_fac;
            v = mk_elem<M>(a, 0.0);
        } else if (!psi) {
            double a;
            if (r == 0)
                a = c >= nc ? 0.0 : (p.ktopv == 1 ? rn * (PZ_CDM(0, c) - R.or1(0) * PZ_CM(0, c)) : rn * PZ_CM(0, c));
            else if (r == N - 1)
                a = c >= nc ? 0.0
                            : (p.kbotv == 1 ? rn * (PZ_CDM(N - 1, c) - R.or1(N - 1) * PZ_CM(N - 1, c))
                                            : rn * PZ_CM(N - 1, c));
            else
                a = rn * (PZ_CM(r, c) - wimp * (p.ViscFac * hdif * (PZ_CD2M(r, c) + R.or1(r) * PZ_CDM(r, c) -
                                                                    R.or2(r) * PZ_CM(r, c)) -
                                                 p.CorFac * R.ekpump(r) * PZ_CM(r, c)));
            if (c == 0 || c == N - 1) a *= p.boundary_fac;
            v = mk_elem<M>(a, 0.0);
        } else {
            // rows r < N: omega equation (BC rows 0, N-1: non-penetration
            // on psi); rows r >= N: omega = Laplacian-type relation (BC rows
            // N, 2N-1: slip condition on psi); columns c < N: omega, c >= N: psi
            const int rr = r < N ? r : r - N, cc = c < N ? c : c - N;
            const bool cpsi = c >= N;
            double re = 0.0, im = 0.0;
            if (rr == 0 || rr == N - 1) {
                if (cpsi && cc < nc) {
                    const int k = rr;
                    if (r < N) re = rn * PZ_CM(k, cc);
                    else {
                        const int kv = r == N ? p.ktopv : p.kbotv;
                        re = kv == 1 ? rn * (PZ_CD2M(k, cc) - R.or1(k) * PZ_CDM(k, cc)) : rn * PZ_CDM(k, cc);
                    }
                }
            } else if (r < N) {
                const double ekp = p.CorFac * R.ekpump(rr);
                if (!cpsi) {
                    re = rn * (PZ_CM(rr, cc) - wimp * (p.ViscFac * hdif * (PZ_CD2M(rr, cc) + R.or1(rr) * PZ_CDM(rr, cc) -
                                                                   dm2 * R.or2(rr) * PZ_CM(rr, cc)) -
                                                       ekp * PZ_CM(rr, cc)));
                } else {
                    // -rn wimp (ekp ekp_up dr + ekp (ekp_up beta - (i m ekp_dusdp + ekp_us) i m / r) r)
                    //  (- rn wimp CorFac beta / r i m r with implicit Coriolis)
                    const double up = R.ekp_up(rr), us = R.ekp_us(rr), dsp = R.ekp_dusdp(rr), be = R.beta(rr);
                    const double o1 = R.or1(rr);
                    // (i m dsp + us) i m / r = (-m^2 dsp + i m us) / r
                    const double gre = up * be + dm2 * dsp * o1, gim = -dm * us * o1;
                    re = -rn * wimp * (ekp * up * PZ_CDM(rr, cc) + ekp * gre * PZ_CM(rr, cc));
                    im = -rn * wimp * (ekp * gim * PZ_CM(rr, cc));
                    if (p.l_coriolis_imp) im -= rn * wimp * p.CorFac * be * o1 * dm * PZ_CM(rr, cc);
                }
            } else {
                if (!cpsi) re = rn * PZ_CM(rr, cc);
                else
                    re = rn * (PZ_CD2M(rr, cc) + (R.or1(rr) + R.beta(rr)) * PZ_CDM(rr, cc) +
                               (R.or1(rr) * R.beta(rr) + R.dbeta(rr) - dm2 * R.or2(rr)) * PZ_CM(rr, cc));
            }
            if (cc == 0 || cc == N - 1) { re *= p.boundary_fac; im *= p.boundary_fac; }
            v = mk_elem<M>(re, im);
        }
        W[(size_t)r * w2 + c] = v;
    }
    __syncthreads();
    if (fld < 2 || psi) {
        // row max-scaling (|.| of the reference: the modulus)
        for (int r = tid; r < n; r += nt) {
            double mx = 0.0;
            for (int c = 0; c < n; c++) mx = fmax(mx, mabs2(W[(size_t)r * w2 + c]));
            fac[r] = 1.0 / mx;
            for (int c = 0; c < n; c++) W[(size_t)r * w2 + c] = mscale(W[(size_t)r * w2 + c], fac[r]);
        }
        __syncthreads();
    }
    if (psi) {
        for (int c = tid; c < n; c += nt) {
            double mx = 0.0;
            for (int r = 0; r < n; r++) mx = fmax(mx, mabs2(W[(size_t)r * w2 + c]));
            fac[n + c] = 1.0 / mx;
            for (int r = 0; r < n; r++) W[(size_t)r * w2 + c] = mscale(W[(size_t)r * w2 + c], fac[n + c]);
        }
        __syncthreads();
    }
}

// get_omMat / get_psiMat of the influence-matrix method
// (update_psi_coll_dmat.f90:829-983) into the left half of W (row-major,
// 2N wide): which = 0 the omega operator (identity minus the implicit
// viscous term), 1 the stream-function operator with beta; zero rows at
// both walls, boundary_fac on the first and last columns, then the row
// (fac[0..N)) and column (fac[N..2N)) max-scalings
__device__ void coll_assemble_dmat(const PzDev &p, const RadView &R, int which, int m, double hdif, double *W,
                                   double *fac)
{
    const int N = p.N, nc = p.ncheb, tid = threadIdx.x, nt = blockDim.x, w2 = 2 * N;
    const double wimp = p.ss->wimp_lin[0], rn = p.rnorm, dm2 = (double)m * (double)m;
    for (int t = tid; t < N * N; t += nt) {
        const int r = t / N, c = t - r * N;
        double a;
        if (r == 0 || r == N - 1)
            a = c < nc ? rn * PZ_CM(r, c) : 0.0;
        else if (which == 0)
            a = rn * (PZ_CM(r, c) -
                      wimp * p.ViscFac * hdif * (PZ_CD2M(r, c) + R.or1(r) * PZ_CDM(r, c) - dm2 * R.or2(r) * PZ_CM(r, c)));
        else
            a = rn * (PZ_CD2M(r, c) + (R.or1(r) + R.beta(r)) * PZ_CDM(r, c) +
                      (R.or1(r) * R.beta(r) + R.dbeta(r) - dm2 * R.or2(r)) * PZ_CM(r, c));
        if (c == 0 || c == N - 1) a *= p.boundary_fac;
        W[(size_t)r * w2 + c] = a;
    }
    __syncthreads();
    for (int r = tid; r < N; r += nt) {
        double mx = 0.0;
        for (int c = 0; c < N; c++) mx = fmax(mx, fabs(W[(size_t)r * w2 + c]));
        fac[r] = 1.0 / mx;
    }
    __syncthreads();
    for (int t = tid; t < N * N; t += nt) {
        const int r = t / N, c = t - r * N;
        W[(size_t)r * w2 + c] *= fac[r];
    }
    __syncthreads();
    for (int c = tid; c < N; c += nt) {
        double mx = 0.0;
        for (int r = 0; r < N; r++) mx = fmax(mx, fabs(W[(size_t)r * w2 + c]));
        fac[N + c] = 1.0 / mx;
    }
    __syncthreads();
    for (int t = tid; t < N * N; t += nt) {
        const int r = t / N, c = t - r * N;
        W[(size_t)r * w2 + c] *= fac[N + c];
    }
    __syncthreads();
}

// grid (n_m, 3): field 0 = T, 1 = xi, 2 = psi (m /= 0) or u_phi (m = 0);
// with the influence-matrix method (n_m, 4): 2 = omega (m /= 0) or u_phi,
// 3 = psi (m /= 0).
// Forms the inverses, with the reference's scalings folded in, whenever the
// matrices change (lMat) or when forced; only the first n_r columns of the
// psi inverse are formed, the ones the m-loop applies.  Dynamic LDS:
// gj_lds_bytes<cd>(2 n_r, p.gj_b).
__device__ __forceinline__ void build_coll_body(const PzDev &p, int mi, int fld, int force)
{
    __shared__ double facs[4 * 257];
    extern __shared__ double2 gj_smem[]; // gj_lds_bytes<cd>(2 n_r, p.gj_b) bytes
    if (!force && p.ss->lMat == 0) return;
    const int N = p.N, m = p.idx2m[mi], tid = threadIdx.x;
    if ((fld == 0 && !p.l_heat) || (fld == 1 && !p.l_chem)) return;
    const RadView R(p.radc, N);
    char *ws = p.coll_work + (size_t)mi * p.coll_ws;
    const size_t regT = (size_t)N * 2 * N * sizeof(double);
    const int ldi = p.ldi;
    if (p.dmat && fld >= 2 && m != 0) {
        // influence-matrix method: field 2 the omega operator, 3 the psi
        // operator; INV = diag(fac_col) A_s^{-1} diag(fac_row), real
        double *W = (double *)(ws + fld * regT);
        coll_assemble_dmat(p, R, fld - 2, m, p.hdifV[mi], W, facs);
        const int info = gj_invert_blk<double>(W, N, N, (double *)gj_smem, p.gj_b);
        if (tid == 0) p.info[fld * p.nm + mi] = info;
        if (info) return;
        double *out = (fld == 2 ? p.OmI : p.PsI) + (size_t)mi * N * ldi;
        for (int t = tid; t < N * ldi; t += blockDim.x) {
            const int c = t / ldi, r = t - c * ldi;
            out[(size_t)c * ldi + r] = r < N ? facs[N + r] * W[(size_t)r * 2 * N + N + c] * facs[c] : 0.0;
        }
        return;
    }
    if (fld == 3) return;
    if (fld < 2 || m == 0) {
        double *W = (double *)(ws + (fld == 2 ? 2 * regT : fld * regT));
        coll_assemble<double>(p, R, fld, m, fld == 0 ? p.hdifT[mi] : fld == 1 ? p.hdifXi[mi] : p.hdifV[mi], W, facs);
        const int info = gj_invert_blk<double>(W, N, N, (double *)gj_smem, p.gj_b);
        if (tid == 0) p.info[fld * p.nm + mi] = info;
        if (info) return;
        // INV = A_s^{-1} diag(fac) (T, xi); A^{-1} (u_phi); column-major [col][ldi]
        double *out = fld == 0 ? p.T_inv + (size_t)mi * N * ldi : fld == 1 ? p.X_inv + (size_t)mi * N * ldi : p.U_inv;
        for (int t = tid; t < N * ldi; t += blockDim.x) {
            const int c = t / ldi, r = t - c * ldi;
            double v = 0.0;
            if (r < N) {
                v = W[(size_t)r * 2 * N + N + c];
                if (fld < 2) v *= facs[c];
            }
            out[(size_t)c * ldi + r] = v;
        }
    } else {
        cd *W = (cd *)(ws + 2 * regT);
        const int n = 2 * N;
        coll_assemble<cd>(p, R, fld, m, p.hdifV[mi], W, facs);
        const int info = gj_invert_blk<cd>(W, n, N, (cd *)gj_smem, p.gj_b);
        if (tid == 0) p.info[fld * p.nm + mi] = info;
        if (info) return;
        // INV = diag(fac_col) A_s^{-1} diag(fac_row); only the first N
        // columns (the omega-equation rows of the right-hand side) are kept
        cd *out = p.PC_inv + (size_t)mi * N * n;
        for (int t = tid; t < N * n; t += blockDim.x) {
            const int c = t / n, r = t - c * n;
            out[(size_t)c * n + r] = cscale(facs[n + r] * facs[c], W[(size_t)r * 2 * n + n + c]);
        }
    }
}

// dynamic LDS of k_dmat_homog: omega_2,3 and their grid values (4 N
// complex), the transform scratch (4 L + 16 complex), the psi right-hand
// sides (2 N doubles)
__host__ __device__ inline size_t dmat_homog_lds(int N, int L)
{
    return (4 * (size_t)N + 4 * (size_t)L + 16) * sizeof(cd) + 2 * (size_t)N * sizeof(double);
}

// The influence-matrix set-up of one m, after the inverses are built
// (update_psi_coll_dmat.f90:182-205, 296-312, 344-347, 1136-1254):
// omega_2 = INV_om e_(N-1) (omega = 1 at the inner wall), omega_3 =
// INV_om e_0 (outer wall), psi_2,3 = INV_psi (0, -omega_2,3 on the grid, 0),
// and the inverse of the 2 x 2 matrix of the wall conditions of psi_2,3.
// One workgroup per m; m = 0 has no influence matrix.
__global__ void __launch_bounds__(256) k_dmat_homog(const PzDev *__restrict__ pd, int force)
{
    const PzDev &p = *pd;
    if (!force && p.ss->lMat == 0) return;
    const int mi = blockIdx.x;
    if (p.idx2m[mi] == 0) return;
    extern __shared__ double smem[];
    const int N = p.N, nc = p.ncheb, tid = threadIdx.x, ldi = p.ldi;
    cd *o = (cd *)smem, *g = o + 2 * N, *buf = g + 2 * N;
    double *rhs = (double *)(buf + 4 * p.L + 16);
    const double *OmI = p.OmI + (size_t)mi * N * ldi, *PsI = p.PsI + (size_t)mi * N * ldi;
    double *hom = p.dhom + (size_t)mi * 4 * N;
    for (int k = tid; k < N; k += blockDim.x) {
        const double a = k < nc ? OmI[(size_t)(N - 1) * ldi + k] : 0.0, b = k < nc ? OmI[k] : 0.0;
        o[k] = mk(a, 0.0);
        o[N + k] = mk(b, 0.0);
        hom[k] = a;
        hom[N + k] = b;
    }
    __syncthreads();
    dct_fft(p, VecSet<const cd>(o, o + N), VecSet<cd>(g, g + N), 2, 0, p.cosT, p.sinT, buf);
    __syncthreads();
    for (int t = tid; t < 2 * N; t += blockDim.x) {
        const int k = t % N;
        rhs[t] = (k == 0 || k == N - 1) ? 0.0 : -g[t].x;
    }
    __syncthreads();
    // psi_2,3 = INV_psi rhs, one row per thread (columns of INV_psi are
    // contiguous, so the rows of a wave read consecutive addresses)
    for (int t = tid; t < 2 * N; t += blockDim.x) {
        const int set = t / N, r = t - set * N;
        double v = 0.0;
        if (r < nc)
            for (int c = 0; c < N; c++) v += PsI[(size_t)c * ldi + r] * rhs[(size_t)set * N + c];
        hom[(2 + set) * (size_t)N + r] = v;
    }
    __syncthreads();
    if (tid == 0) {
        double t2 = 0, b2 = 0, t3 = 0, b3 = 0;
        for (int k = 0; k < N; k++) {
            t2 += p.bcrow[k] * hom[2 * (size_t)N + k];
            b2 += p.bcrow[N + k] * hom[2 * (size_t)N + k];
            t3 += p.bcrow[k] * hom[3 * (size_t)N + k];
            b3 += p.bcrow[N + k] * hom[3 * (size_t)N + k];
        }
        const double det = t2 * b3 - t3 * b2;
        double *inf = p.dinfl + (size_t)mi * 4;
        inf[0] = b3 / det;
        inf[1] = -t3 / det;
        inf[2] = -b2 / det;
        inf[3] = t2 / det;
    }
}

__global__ void __launch_bounds__(1024) k_build_coll(const PzDev *__restrict__ pd, int force)
{
    build_coll_body(*pd, blockIdx.x, blockIdx.y, force);
}
// batched: local block index = mi + nm fld; a grid-stride loop over the
// nb blocks of all runs, so that the usual early exit costs few workgroups
__global__ void __launch_bounds__(1024) k_build_coll_batch(PzBatch b, int nb, int force)
{
    if (!batch_any_lmat(b, force)) return;
    for (int i = blockIdx.x; i < nb; i += gridDim.x) {
        const int2 e = b.map[i];
        const PzDev &p = *b.runs[e.x];
        build_coll_body(p, e.y % p.nm, e.y / p.nm, force);
        __syncthreads();
    }
}

// get_ddcheb_real_1d on Chebyshev coefficients f (radial_derivatives.f90:
// 255-311), one thread; df, ddf are coefficients
__device__ __forceinline__ void ddcheb_real(const PzDev &p, const double *__restrict__ f, double *__restrict__ df,
                                            double *__restrict__ ddf)
{
    const int N = p.N;
    // four independent maxima so that the loads overlap
    double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
    int j = 0;
    for (; j + 3 < N; j += 4) {
        m0 = fmax(m0, fabs(f[j]));
        m1 = fmax(m1, fabs(f[j + 1]));
        m2 = fmax(m2, fabs(f[j + 2]));
        m3 = fmax(m3, fabs(f[j + 3]));
    }
    for (; j < N; j++) m0 = fmax(m0, fabs(f[j]));
    const double thr = fmax(fmax(m0, m1), fmax(m2, m3)) * p.thr;
    // both recurrences carried in registers: d1 = df(n), d2 = df(n+1),
    // e1 = ddf(n), e2 = ddf(n+1)
    df[N - 1] = 0.0;
    ddf[N - 1] = 0.0;
    const int nc = N - 1;
    double d1 = fabs(f[nc]) >= thr ? (double)nc * f[nc] : 0.0, d2 = 0.0, e1 = 0.0, e2 = 0.0;
    df[nc - 1] = d1;
    ddf[nc - 1] = 0.0;
#pragma unroll 4
    for (int n = N - 2; n >= 1; n--) {
        const double fac = (double)(2 * n), fn = f[n];
        const double dv = fabs(fn) >= thr ? d2 + fac * fn : d2;
        const double ev = fabs(d1) >= thr ? e2 + fac * d1 : e2;
        df[n - 1] = dv;
        ddf[n - 1] = ev;
        d2 = d1;
        d1 = dv;
        e2 = e1;
        e1 = ev;
    }
}

// m = 0 implicit term of get_psi_rhs_imp_coll_smat (update_psi_coll_smat.f90:
// 454-466): u0 holds u_phi(m = 0) on the grid (real parts); writes impl.
// Uses cA, cB (LDS vectors) and dt3 (3 N doubles) as scratch.
__device__ __forceinline__ void coll_uphi0_impl(const PzDev &p, const RadView &R, const cd *u0, cd *cA, cd *cB, double *dt3,
                                const double *cosT, const double *sinT, cd *fbuf, cd *impl_out, double hdif)
{
    const int N = p.N, tid = threadIdx.x;
    // get_ddr_real_1d (radial_derivatives.f90:593-667): costf, recursion, costf back, map
    {
        const VecSet<const cd> din(u0);
        VecSet<cd> dout(cA);
        dct_fft(p, din, dout, 1, 0, cosT, sinT, fbuf);
    }
    __syncthreads();
    double *f = dt3, *df = dt3 + N, *ddf = dt3 + 2 * N;
    for (int k = tid; k < N; k += blockDim.x) f[k] = cA[k].x;
    __syncthreads();
    if (tid == 0) ddcheb_real(p, f, df, ddf);
    __syncthreads();
    for (int k = tid; k < N; k += blockDim.x) { cA[k] = mk(df[k], 0.0); cB[k] = mk(ddf[k], 0.0); }
    __syncthreads();
    {
        const VecSet<const cd> din(cA, cB);
        VecSet<cd> dout(cA, cB);
        dct_fft(p, din, dout, 2, 0, cosT, sinT, fbuf);
    }
    __syncthreads();
    for (int k = tid; k < N; k += blockDim.x) {
        const double a = R.drx(k), b = R.ddrx(k);
        const double d1 = a * cA[k].x, d2 = b * cA[k].x + a * a * cB[k].x, u = u0[k].x;
        const double v = p.ViscFac * hdif * (d2 + R.or1(k) * d1 - R.or2(k) * u) - p.CorFac * R.ekpump(k) * u;
        impl_out[k] = mk(v, 0.0);
    }
}

// One stage of the m-loop, collocation method, one azimuthal mode per
// workgroup: finish_exp_*_coll (update_temp_coll.f90:152-176,
// update_xi_coll.f90:148-166, update_psi_coll_smat.f90:323-378) when
// exp_calc, update_temp_co / update_xi_co (update_temp_coll.f90:66-150) and
// update_om_coll_smat (update_psi_coll_smat.f90:88-321) with their implicit
// terms (get_*_rhs_imp_coll*).
template <bool RADLDS>
__device__ __forceinline__ void mcoll_body(const PzDev &p, int mi, int istage, int exp_calc)
{
    extern __shared__ double smem[];
    const int N = p.N, m = p.idx2m[mi];
    const size_t F = (size_t)p.nm * N, off = (size_t)mi * N;
    const int is = istage - 1;
    const PzScheme *ss = p.ss;
    const int tid = threadIdx.x;
    const MpLayout lay(N, blockDim.x, RADLDS);
    char *lb = (char *)smem;
    cd *V = (cd *)(lb + lay.vec);
    cd *aT = V, *cT = V + N, *eT = V + 2 * N;
    cd *aX = V + 3 * N, *cX = V + 4 * N, *eX = V + 5 * N;
    cd *aP = V + 6 * N, *cP = V + 7 * N, *eP = V + 8 * N; // aP and cP together hold 2N values
    cd *psiv = V + 9 * N, *omv = V + 10 * N, *upv = V + 11 * N, *usv = V + 12 * N;
    cd *Tv = V + 13 * N, *Xv = V + 14 * N, *bv = V + 15 * N, *b2 = V + 16 * N, *wv = V + 17 * N;
    cd *PT = V + 18 * N, *QT = V + 19 * N, *PX = V + 20 * N, *QX = V + 21 * N, *PP = V + 22 * N, *QP = V + 23 * N;
    cd *red = (cd *)(lb + lay.red), *fbuf = red;
    const bool hT = p.l_heat, hX = p.l_chem;
    const int ldi = p.ldi, nc = p.ncheb;
    const double dm = (double)m, dm2 = dm * dm;
    // transform tables and radial rows: LDS copies, or global (L2) reads
    const double *cosT = p.cosT, *sinT = p.sinT;
    if (RADLDS) {
        double *c = (double *)(lb + lay.cosT), *s = (double *)(lb + lay.sinT);
        lds_copy(c, p.cosT, sizeof(double) * p.L);
        lds_copy(s, p.sinT, sizeof(double) * p.L);
        cosT = c;
        sinT = s;
        lds_copy(lb + lay.rad, p.radc, sizeof(double) * PZ_NRADC * N);
    }
    const RadView R(RADLDS ? (const double *)(lb + lay.rad) : p.radc, N);

    // ---------------- global inputs -------------------------------------
    // three groups of N threads: T, xi, psi (three times as many loads in
    // flight)
    for (int t = tid; t < 3 * N; t += blockDim.x) {
        const int g = t / N, k = t - g * N;
        if (g == 0) {
            if (hT) eT[k] = p.T_expl[is * F + off + k];
            Tv[k] = hT ? p.temp[off + k] : mk(0, 0);
            if (exp_calc && hT) aT[k] = p.dVsT[off + k];
            if (hT) imex_split(p, ss, istage, p.T_old, p.T_impl, p.T_expl, off, k, PT[k], QT[k]);
        } else if (g == 1) {
            if (hX) eX[k] = p.X_expl[is * F + off + k];
            Xv[k] = hX ? p.xi[off + k] : mk(0, 0);
            if (exp_calc && hX) aX[k] = p.dVsXi[off + k];
            if (hX) imex_split(p, ss, istage, p.X_old, p.X_impl, p.X_expl, off, k, PX[k], QX[k]);
        } else {
            eP[k] = p.P_expl[is * F + off + k];
            usv[k] = p.us[off + k];
            if (exp_calc) aP[k] = p.dVsOm[off + k];
            if (exp_calc && p.dmat && p.l_ek_pump) {
                omv[k] = p.om[off + k];
                upv[k] = p.up[off + k];
            }
            imex_split(p, ss, istage, p.P_old, p.P_impl, p.P_expl, off, k, PP[k], QP[k]);
        }
    }
    const double wcur = imex_weight(p, ss, istage);
    // the groups above wrote vectors that other threads read next, with or
    // without the explicit part
    __syncthreads();

    // ---------------- explicit part --------------------------------------
    if (exp_calc) {
        {
            VecSet<const cd> din;
            VecSet<cd> dout;
            if (hT) { din.push(aT); dout.push(cT); }
            if (hX) { din.push(aX); dout.push(cX); }
            din.push(aP); dout.push(cP);
            dct_fft(p, din, dout, din.n, 0, cosT, sinT, fbuf);
        }
        __syncthreads();
        {
            VecSet<const cd> c3;
            VecSet<cd> d3;
            if (hT) { c3.push(cT); d3.push(aT); }
            if (hX) { c3.push(cX); d3.push(aX); }
            c3.push(cP); d3.push(aP);
            dr_fft(p, R, c3, d3, VecSet<cd>(), c3.n, cosT, sinT, fbuf);
        }
        __syncthreads();
        for (int k = tid; k < N; k += blockDim.x) {
            const double o1 = R.or1(k);
            if (hT) {
                eT[k] = csub(csub(eT[k], cscale(o1, aT[k])),
                             cscale(R.dtcond(k) - p.tadvz * R.beta(k) * R.tcond(k), usv[k]));
                p.T_expl[is * F + off + k] = eT[k];
            }
            if (hX) {
                eX[k] = csub(csub(eX[k], cscale(o1, aX[k])), cscale(R.dxicond(k), usv[k]));
                p.X_expl[is * F + off + k] = eX[k];
            }
            if (m != 0) {
                cd e = csub(eP[k], cscale(o1, aP[k]));
                if (!p.l_coriolis_imp) e = cadd(e, cscale(p.CorFac * R.beta(k), usv[k]));
                if (!p.l_buo_imp) {
                    if (hT) e = csub(e, cmul(mk(0.0, p.BuoFac * R.rgrav(k) * o1 * dm), Tv[k]));
                    if (hX) e = csub(e, cmul(mk(0.0, p.ChemFac * R.rgrav(k) * o1 * dm), Xv[k]));
                }
                if (p.dmat && p.l_ek_pump) {
                    // finish_exp_psi_coll_dmat (update_psi_coll_dmat.f90:527-549):
                    // pumping explicit with the influence-matrix method
                    cd pump = cadd(omv[k], cscale(R.ekp_up(k), upv[k]));
                    pump = cadd(pump, cmul(mk(R.ekp_us(k), dm * R.ekp_dusdp(k)), usv[k]));
                    e = csub(e, cscale(p.CorFac * R.ekpump(k), pump));
                }
                eP[k] = e;
                p.P_expl[is * F + off + k] = e;
            }
        }
    }

    // ---------------- T and xi -------------------------------------------
    for (int k = tid; k < N; k += blockDim.x) {
        if (hT) {
            cd r = cadd(cadd(PT[k], cscale(wcur, eT[k])), QT[k]);
            if (k == 0) r = p.topt[mi];
            if (k == N - 1) r = p.bott[mi];
            aT[k] = r;
        }
        if (hX) {
            cd r = cadd(cadd(PX[k], cscale(wcur, eX[k])), QX[k]);
            if (k == 0) r = p.topxi[mi];
            if (k == N - 1) r = p.botxi[mi];
            aX[k] = r;
        }
    }
    __syncthreads();
    if (hT && hX) {
        const VecSet<const double> inv(p.T_inv + off * ldi, p.X_inv + off * ldi);
        const VecSet<cd> xs(aT, aX);
        block_matvec<double, 2>(N, ldi, inv, xs, xs, red);
    } else if (hT || hX) {
        const VecSet<const double> inv(hT ? p.T_inv + off * ldi : p.X_inv + off * ldi);
        const VecSet<cd> xs(hT ? aT : aX);
        block_matvec<double, 1>(N, ldi, inv, xs, xs, red);
    }
    for (int k = nc + tid; k < N; k += blockDim.x) { aT[k] = mk(0, 0); aX[k] = mk(0, 0); }
    __syncthreads();
    const int inext = (istage == p.nstages) ? 1 : istage + 1;
    // implicit terms for the next stage; also before a CNAB2 start-up step
    // of a multistep scheme, which recomputes them (step_time.f90:417-503)
    const int lcalc = p.l_imp_calc_rhs[inext - 1] || (p.bridge && ss->n_steps + 1 <= p.nold - 1);
    if (lcalc) {
        // get_ddr (coefficients in) and costf: derivatives and grid values
        if (hT) {
            dr_fft(p, R, VecSet<const cd>(aT), VecSet<cd>(cT), VecSet<cd>(bv), 1, cosT, sinT, fbuf, Tv);
            __syncthreads();
        }
        if (hX) {
            dr_fft(p, R, VecSet<const cd>(aX), VecSet<cd>(cX), VecSet<cd>(b2), 1, cosT, sinT, fbuf, Xv);
            __syncthreads();
        }
    } else {
        VecSet<const cd> din;
        VecSet<cd> dout;
        if (hT) { din.push(aT); dout.push(Tv); }
        if (hX) { din.push(aX); dout.push(Xv); }
        if (din.n) dct_fft(p, din, dout, din.n, 0, cosT, sinT, fbuf);
        __syncthreads();
    }
    if (hT) rotate_row(p, p.T_old, p.T_impl, p.T_expl, off);
    if (hX) rotate_row(p, p.X_old, p.X_impl, p.X_expl, off);
    __syncthreads();
    for (int k = tid; k < N; k += blockDim.x) {
        const double o1 = R.or1(k), o2 = R.or2(k);
        if (hT) {
            p.temp[off + k] = Tv[k];
            if (inext == 1) p.T_old[off + k] = Tv[k];
            if (lcalc) {
                p.dtemp[off + k] = cT[k];
                const cd v = csub(cadd(bv[k], cscale(o1, cT[k])), cscale(dm2 * o2, Tv[k]));
                p.T_impl[(inext - 1) * F + off + k] = cscale(p.TdiffFac * p.hdifT[mi], v);
            }
        }
        if (hX) {
            p.xi[off + k] = Xv[k];
            if (inext == 1) p.X_old[off + k] = Xv[k];
            if (lcalc) {
                const cd v = csub(cadd(b2[k], cscale(o1, cX[k])), cscale(dm2 * o2, Xv[k]));
                p.X_impl[(inext - 1) * F + off + k] = cscale(p.XidiffFac * p.hdifXi[mi], v);
            }
        }
    }
    __syncthreads();

    // ---------------- psi / omega / u_phi ---------------------------------
    const double wl = ss->wimp_lin[0];
    if (m == 0) {
        // u_phi(m = 0): real system, zero boundary rows
        for (int k = tid; k < N; k += blockDim.x) {
            const cd r = cadd(cadd(PP[k], cscale(wcur, eP[k])), QP[k]);
            wv[k] = (k == 0 || k == N - 1) ? mk(0, 0) : mk(r.x, 0.0);
        }
        __syncthreads();
        block_matvec<double, 1>(N, ldi, VecSet<const double>(p.U_inv), VecSet<cd>(wv), VecSet<cd>(wv), red);
        for (int k = nc + tid; k < N; k += blockDim.x) wv[k] = mk(0, 0);
        __syncthreads();
        // om0 = get_dr(uphi0, l_dct = .false.) (dcheb recursion), uphi0 -> grid
        double *f = (double *)b2, *df = f + N;
        for (int k = tid; k < N; k += blockDim.x) f[k] = wv[k].x;
        __syncthreads();
        if (tid == 0) dcheb_real(p, f, df);
        __syncthreads();
        for (int k = tid; k < N; k += blockDim.x) { cP[k] = mk(df[k], 0.0); eP[k] = mk(f[k], 0.0); }
        __syncthreads();
        {
            const VecSet<const cd> din(cP, eP);
            VecSet<cd> dout(omv, upv);
            dct_fft(p, din, dout, 2, 0, cosT, sinT, fbuf);
        }
        __syncthreads();
        for (int k = tid; k < N; k += blockDim.x) {
            const double u0 = upv[k].x, om0 = R.drx(k) * omv[k].x;
            const cd om = mk(om0 + R.or1(k) * u0, 0.0);
            upv[k] = mk(u0, 0.0);
            omv[k] = om;
            p.us[off + k] = mk(0, 0);
            p.up[off + k] = upv[k];
            p.om[off + k] = om;
            p.psi[off + k] = mk(0, 0);
        }
        rotate_row(p, p.P_old, p.P_impl, p.P_expl, off);
        __syncthreads();
        if (inext == 1)
            for (int k = tid; k < N; k += blockDim.x) p.P_old[off + k] = upv[k];
        if (lcalc)
            coll_uphi0_impl(p, R, upv, cP, eP, (double *)bv, cosT, sinT, fbuf, wv, p.hdifV[mi]);
        __syncthreads();
        if (lcalc)
            for (int k = tid; k < N; k += blockDim.x) p.P_impl[(inext - 1) * F + off + k] = wv[k];
        return;
    }
    for (int k = tid; k < N; k += blockDim.x) {
        cd r = cadd(cadd(PP[k], cscale(wcur, eP[k])), QP[k]);
        if (p.l_buo_imp) {
            if (hT) r = csub(r, cmul(mk(0.0, wl * p.BuoFac * R.rgrav(k) * R.or1(k) * dm), Tv[k]));
            if (hX) r = csub(r, cmul(mk(0.0, wl * p.ChemFac * R.rgrav(k) * R.or1(k) * dm), Xv[k]));
        }
        aP[k] = (k == 0 || k == N - 1) ? mk(0, 0) : r;
    }
    __syncthreads();
    if (p.dmat) {
        // update_om_coll_dmat (update_psi_coll_dmat.f90:244-453): omega_1
        // from its own equation (omega = 0 at the walls) ...
        block_matvec<double, 1>(N, ldi, VecSet<const double>(p.OmI + off * ldi), VecSet<cd>(aP), VecSet<cd>(aP), red);
        for (int k = nc + tid; k < N; k += blockDim.x) aP[k] = mk(0, 0);
        __syncthreads();
        dct_fft(p, VecSet<const cd>(aP), VecSet<cd>(eP), 1, 0, cosT, sinT, fbuf);
        __syncthreads();
        // ... psi_1 from L_psi psi = -omega_1 (psi = 0 at the walls) ...
        for (int k = tid; k < N; k += blockDim.x) cP[k] = (k == 0 || k == N - 1) ? mk(0, 0) : mk(-eP[k].x, -eP[k].y);
        __syncthreads();
        block_matvec<double, 1>(N, ldi, VecSet<const double>(p.PsI + off * ldi), VecSet<cd>(cP), VecSet<cd>(cP), red);
        for (int k = nc + tid; k < N; k += blockDim.x) cP[k] = mk(0, 0);
        __syncthreads();
        // ... and the multiples of the homogeneous pairs that meet the
        // velocity conditions (solve_influence_matrix, :1256-1300)
        double tx = 0.0, ty = 0.0, bx = 0.0, by = 0.0;
        for (int k = tid; k < N; k += blockDim.x) {
            const double ct = p.bcrow[k], cb = p.bcrow[N + k];
            tx += ct * cP[k].x;
            ty += ct * cP[k].y;
            bx += cb * cP[k].x;
            by += cb * cP[k].y;
        }
        tx = block_reduce(tx, (double *)red, 0);
        ty = block_reduce(ty, (double *)red, 0);
        bx = block_reduce(bx, (double *)red, 0);
        by = block_reduce(by, (double *)red, 0);
        const double *inf = p.dinfl + (size_t)mi * 4, *hom = p.dhom + (size_t)mi * 4 * N;
        const cd c1 = mk(-inf[0] * tx - inf[1] * bx, -inf[0] * ty - inf[1] * by);
        const cd c2 = mk(-inf[2] * tx - inf[3] * bx, -inf[2] * ty - inf[3] * by);
        for (int k = tid; k < N; k += blockDim.x) {
            aP[k] = cadd(cadd(aP[k], cscale(hom[k], c1)), cscale(hom[N + k], c2));
            cP[k] = cadd(cadd(cP[k], cscale(hom[2 * N + k], c1)), cscale(hom[3 * N + k], c2));
        }
        __syncthreads();
    } else {
        // [omega; psi] coefficients (2N values in aP, cP)
        block_matvec_rc<cd, 1>(2 * N, N, 2 * N, VecSet<const cd>(p.PC_inv + off * 2 * N), VecSet<cd>(aP),
                               VecSet<cd>(aP), red);
        for (int k = nc + tid; k < N; k += blockDim.x) { aP[k] = mk(0, 0); cP[k] = mk(0, 0); }
        __syncthreads();
    }
    // d psi / dr and psi on the grid
    dr_fft(p, R, VecSet<const cd>(cP), VecSet<cd>(eP), VecSet<cd>(), 1, cosT, sinT, fbuf, psiv);
    __syncthreads();
    // omega on the grid, with its derivatives for the implicit term
    if (lcalc) dr_fft(p, R, VecSet<const cd>(aP), VecSet<cd>(bv), VecSet<cd>(b2), 1, cosT, sinT, fbuf, omv);
    else dct_fft(p, VecSet<const cd>(aP), VecSet<cd>(omv), 1, 0, cosT, sinT, fbuf);
    __syncthreads();
    for (int k = tid; k < N; k += blockDim.x) {
        const cd ps = psiv[k];
        const cd us = cmul(mk(0.0, dm * R.or1(k)), ps);
        const cd up = csub(mk(-eP[k].x, -eP[k].y), cscale(R.beta(k), ps));
        usv[k] = us;
        upv[k] = up;
        p.psi[off + k] = ps;
        p.us[off + k] = us;
        p.up[off + k] = up;
        p.om[off + k] = omv[k];
    }
    rotate_row(p, p.P_old, p.P_impl, p.P_expl, off);
    __syncthreads();
    for (int k = tid; k < N; k += blockDim.x) {
        if (inext == 1) p.P_old[off + k] = omv[k];
        if (lcalc) {
            const double o1 = R.or1(k), ekp = p.CorFac * R.ekpump(k);
            cd v = cscale(p.ViscFac * p.hdifV[mi],
                          csub(cadd(b2[k], cscale(o1, bv[k])), cscale(dm2 * R.or2(k), omv[k])));
            if (!p.dmat) {
                // get_psi_rhs_imp_coll_dmat has no pumping term (update_psi_coll_dmat.f90:641-667)
                cd pump = cadd(omv[k], cscale(R.ekp_up(k), upv[k]));
                pump = cadd(pump, cmul(mk(R.ekp_us(k), dm * R.ekp_dusdp(k)), usv[k]));
                v = csub(v, cscale(ekp, pump));
            }
            if (p.l_coriolis_imp) v = cadd(v, cscale(p.CorFac * R.beta(k), usv[k]));
            if (p.l_buo_imp) {
                if (hT) v = csub(v, cmul(mk(0.0, R.rgrav(k) * o1 * p.BuoFac * dm), Tv[k]));
                if (hX) v = csub(v, cmul(mk(0.0, R.rgrav(k) * o1 * p.ChemFac * dm), Xv[k]));
            }
            p.P_impl[(inext - 1) * F + off + k] = v;
        }
    }
}

__global__ void __launch_bounds__(PZ_MPHASE_THREADS) k_mcoll(const PzDev *__restrict__ pd, int istage, int exp_calc)
{
    mcoll_body<true>(*pd, blockIdx.x, istage, exp_calc);
}
// the same with the radial rows and transform tables read through L2, for
// grids whose full LDS layout does not fit a workgroup (n_r > 193)
__global__ void __launch_bounds__(PZ_MPHASE_THREADS) k_mcoll_l2(const PzDev *__restrict__ pd, int istage, int exp_calc)
{
    mcoll_body<false>(*pd, blockIdx.x, istage, exp_calc);
}
// batched; RADLDS picks the radial rows' place per size class
template <bool RADLDS>
__global__ void __launch_bounds__(PZ_MPHASE_THREADS) k_mcoll_batch(PzBatch b, int istage, int exp_calc)
{
    const int2 e = b.map[blockIdx.x];
    mcoll_body<RADLDS>(*b.runs[e.x], e.y, istage, exp_calc);
}

// get_start_fields, collocation branch (init_fields.f90:191-244): dT/dr,
// omega from u_phi and u_s, and the first implicit state from grid values.
// the descriptor comes through a pointer: the out-of-line transforms take
// its address, which a by-value kernel argument would copy to scratch
__global__ void __launch_bounds__(256) k_start_coll(const PzDev *__restrict__ pd)
{
    const PzDev &p = *pd;
    extern __shared__ double smem[];
    const int mi = blockIdx.x, N = p.N, m = p.idx2m[mi], tid = threadIdx.x;
    const size_t off = (size_t)mi * N;
    const double dm = (double)m, dm2 = dm * dm;
    const RadView R(p.radc, N);
    cd *V = (cd *)smem;
    cd *Tv = V, *Xv = V + N, *upv = V + 2 * N, *usv = V + 3 * N, *omv = V + 4 * N;
    cd *c1 = V + 5 * N, *c2 = V + 6 * N, *c3 = V + 7 * N, *d1 = V + 8 * N, *d2 = V + 9 * N;
    cd *e1 = V + 10 * N, *e2 = V + 11 * N, *e3 = V + 12 * N;
    cd *fbuf = V + 13 * N; // 6 L + 3 N + 16
    double *dt3 = (double *)(fbuf + 6 * p.L + 3 * N + 16);
    const bool hT = p.l_heat, hX = p.l_chem;
    for (int k = tid; k < N; k += blockDim.x) {
        Tv[k] = hT ? p.temp[off + k] : mk(0, 0);
        Xv[k] = hX ? p.xi[off + k] : mk(0, 0);
        upv[k] = p.up[off + k];
        usv[k] = p.us[off + k];
    }
    __syncthreads();
    // coefficients of T, xi, u_phi
    dct_fft(p, VecSet<const cd>(Tv, Xv, upv), VecSet<cd>(c1, c2, c3), 3, 0, p.cosT, p.sinT, fbuf);
    __syncthreads();
    // omega = du_phi/dr + u_phi/r - i m u_s/r
    dr_fft(p, R, VecSet<const cd>(c3), VecSet<cd>(e3), VecSet<cd>(), 1, p.cosT, p.sinT, fbuf);
    __syncthreads();
    for (int k = tid; k < N; k += blockDim.x) {
        const cd o = cadd(e3[k], cscale(R.or1(k), upv[k]));
        omv[k] = csub(o, cmul(mk(0.0, R.or1(k) * dm), usv[k]));
        p.om[off + k] = omv[k];
    }
    __syncthreads();
    // T and xi: dT/dr, old = T, impl = kappa (T'' + T'/r - m^2 T/r^2)
    if (hT) {
        dr_fft(p, R, VecSet<const cd>(c1), VecSet<cd>(d1), VecSet<cd>(d2), 1, p.cosT, p.sinT, fbuf);
        __syncthreads();
        for (int k = tid; k < N; k += blockDim.x) {
            p.dtemp[off + k] = d1[k];
            p.T_old[off + k] = Tv[k];
            const cd v = csub(cadd(d2[k], cscale(R.or1(k), d1[k])), cscale(dm2 * R.or2(k), Tv[k]));
            p.T_impl[off + k] = cscale(p.TdiffFac * p.hdifT[mi], v);
        }
        __syncthreads();
    }
    if (hX) {
        dr_fft(p, R, VecSet<const cd>(c2), VecSet<cd>(d1), VecSet<cd>(d2), 1, p.cosT, p.sinT, fbuf);
        __syncthreads();
        for (int k = tid; k < N; k += blockDim.x) {
            p.X_old[off + k] = Xv[k];
            const cd v = csub(cadd(d2[k], cscale(R.or1(k), d1[k])), cscale(dm2 * R.or2(k), Xv[k]));
            p.X_impl[off + k] = cscale(p.XidiffFac * p.hdifXi[mi], v);
        }
        __syncthreads();
    }
    // psi equation: old = u_phi (m = 0) or omega, implicit term from grid omega
    if (m == 0) {
        for (int k = tid; k < N; k += blockDim.x) p.P_old[off + k] = upv[k];
        __syncthreads();
        coll_uphi0_impl(p, R, upv, e1, e2, dt3, p.cosT, p.sinT, fbuf, e3, p.hdifV[mi]);
        __syncthreads();
        for (int k = tid; k < N; k += blockDim.x) p.P_impl[off + k] = e3[k];
        return;
    }
    dct_fft(p, VecSet<const cd>(omv), VecSet<cd>(e1), 1, 0, p.cosT, p.sinT, fbuf);
    __syncthreads();
    dr_fft(p, R, VecSet<const cd>(e1), VecSet<cd>(d1), VecSet<cd>(d2), 1, p.cosT, p.sinT, fbuf);
    __syncthreads();
    for (int k = tid; k < N; k += blockDim.x) {
        p.P_old[off + k] = omv[k];
        const double o1 = R.or1(k), ekp = p.CorFac * R.ekpump(k);
        cd v = cscale(p.ViscFac * p.hdifV[mi], csub(cadd(d2[k], cscale(o1, d1[k])), cscale(dm2 * R.or2(k), omv[k])));
        if (!p.dmat) { // the influence-matrix method has no implicit pumping term
            cd pump = cadd(omv[k], cscale(R.ekp_up(k), upv[k]));
            pump = cadd(pump, cmul(mk(R.ekp_us(k), dm * R.ekp_dusdp(k)), usv[k]));
            v = csub(v, cscale(ekp, pump));
        }
        if (p.l_coriolis_imp) v = cadd(v, cscale(p.CorFac * R.beta(k), usv[k]));
        if (p.l_buo_imp) {
            if (hT) v = csub(v, cmul(mk(0.0, R.rgrav(k) * o1 * p.BuoFac * dm), Tv[k]));
            if (hX) v = csub(v, cmul(mk(0.0, R.rgrav(k) * o1 * p.ChemFac * dm), Xv[k]));
        }
        p.P_impl[off + k] = v;
    }
}
