// Oracle: sparse linear algebra of the PFEM-2 path (TEST INFRASTRUCTURE).
// Restates src/linear_algebra/{linear_solver.cu,preconditioners.cu}, src/Dirichlet_bcs.cu and the
// cuBLAS/cuSPARSE calls of src/linear_algebra/linear_algebra.h by their textbook definitions.
#include "oracle.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

// (the OpenMP build reduces in thread chunks; the parity build sums left to right)
double dot(int n, const double* a, const double* b)
{
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static) if (n > 8192)
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

// Diagonal position by the reference's per-row binary search (src/common/cuda_helper.cuh:115-139);
// with ascending columns any search returns the same slot.
int find_col(const int* rowptr, const int* col, int r, int c)
{
    int lo = rowptr[r], hi = rowptr[r + 1] - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        if (col[mid] == c) return mid;
        if (col[mid] < c) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

struct Precond {
    int kind = 0;
    int n = 0;
    const int* rowptr = nullptr;
    const int* col = nullptr;
    std::vector<double> d;   // invD (Jacobi) or factor values (IC0/ILU0)
    bool ok = true;
    void setup(int kind_, int n_, const int* rp, const int* cl, const double* val)
    {
        kind = kind_; n = n_; rowptr = rp; col = cl;
        if (kind == 1) { d.resize(n); orc_jacobi_setup(n, rp, cl, val, d.data()); }
        else if (kind == 2) { d.resize(rp[n]); ok = orc_ic0(n, rp, cl, val, d.data()) == 0; }
        else if (kind == 3) { d.resize(rp[n]); ok = orc_ilu0(n, rp, cl, val, d.data()) == 0; }
    }
    // applyPreconditioner(dest, src); dest may alias src (src/linear_algebra/linear_solver.cu:444,474)
    void apply(double* z, const double* r) const
    {
        if (kind == 0) { if (z != r) std::memcpy(z, r, sizeof(double) * n); }
        else if (kind == 1) {
#pragma omp parallel for schedule(static) if (n > 8192)
            for (int i = 0; i < n; ++i) z[i] = r[i] * d[i];
        }
        else if (kind == 2) {
            std::vector<double> tmp(r, r + n);
            orc_ic0_apply(n, rowptr, col, d.data(), tmp.data(), z);
        } else {
            std::vector<double> tmp(r, r + n);
            orc_ilu0_apply(n, rowptr, col, d.data(), tmp.data(), z);
        }
    }
};

}  // namespace

extern "C" {

// cusparseSpMV CSR (src/linear_algebra/linear_algebra.h:75-80), alpha = 1, beta = 0. The per-row sum
// runs left to right from 0.0 (the order cuSPARSE uses is not specified; the GPU library uses this one).
void orc_spmv(int n, const int* rowptr, const int* col, const double* val, const double* x, double* y)
{
#pragma omp parallel for schedule(static) if (n > 8192)
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) s += val[k] * x[col[k]];
        y[i] = s;
    }
}

// src/Dirichlet_bcs.cu:3-17 (kEnforceValueForBoundaryDoFs) then :19-51 (kAccountForBoundaryValues).
// mode 0 ("parity"): a row c is skipped when |first off-diagonal of c| < 2e-6 (:39-41), evaluated on the
// matrix after enforcement and before any elimination in row c. mode 1 ("mask"): the rows skipped are
// exactly the Dirichlet rows. Elimination order: rows ascending, columns ascending (the reference order
// is that of device atomics and so unspecified; for a node-sorted BC list both coincide).
void orc_dirichlet_apply(int n, const int* rowptr, const int* col, double* val, double* rhs,
                         int nbc, const int* bc_node, const double* bc_val, int mode)
{
    std::vector<int> map(n, -1);
    for (int k = 0; k < nbc; ++k) map[bc_node[k]] = k;
    for (int k = 0; k < nbc; ++k) {
        const int b = bc_node[k];
        for (int e = rowptr[b]; e < rowptr[b + 1]; ++e) val[e] = (col[e] == b) ? 1.0 : 0.0;
        rhs[b] = bc_val[k];
    }
    for (int c = 0; c < n; ++c) {
        const int s = rowptr[c], e = rowptr[c + 1];
        if (mode == 0) {
            const int first = s + (col[s] == c ? 1 : 0);
            if (std::fabs(val[first]) < 2e-6) continue;
        } else if (map[c] >= 0) {
            continue;
        }
        for (int k = s; k < e; ++k) {
            const int b = col[k];
            if (b == c || map[b] < 0) continue;
            const double minus_g = -bc_val[map[b]];
            rhs[c] += val[k] * minus_g;
            val[k] = 0.0;
        }
    }
}

// src/linear_algebra/preconditioners.cu:3-14 (extractDiagonal): invD = 1 / a_rr (no zero guard).
void orc_jacobi_setup(int n, const int* rowptr, const int* col, const double* val, double* invd)
{
    for (int r = 0; r < n; ++r) {
        const int k = find_col(rowptr, col, r, r);
        if (k >= 0) invd[r] = 1.0 / val[k];
    }
}

// Zero-fill incomplete Cholesky on the lower triangle of A (the factor cusparseDcsric02 computes,
// src/linear_algebra/preconditioners.cu:136-213). Row-by-row (left-looking) with every partial sum taken
// in ascending column order. lval has the layout of val; only entries with col <= row are written.
// Returns -(row+1) on a non-positive pivot.
int orc_ic0(int n, const int* rowptr, const int* col, const double* val, double* lval)
{
    for (int i = 0; i < n; ++i) {
        const int si = rowptr[i], ei = rowptr[i + 1];
        int di = -1;
        for (int p = si; p < ei; ++p) {
            const int k = col[p];
            if (k > i) break;
            if (k == i) { di = p; break; }
            // L_ik = (a_ik - sum_{j<k} L_ij L_kj) / L_kk
            double s = val[p];
            int q = rowptr[k], pk = rowptr[k + 1];
            int dk = -1;
            for (int pp = si; pp < p; ++pp) {
                const int j = col[pp];
                while (q < pk && col[q] < j) ++q;
                if (q < pk && col[q] == j) s -= lval[pp] * lval[q];
            }
            for (int qq = rowptr[k]; qq < pk; ++qq) if (col[qq] == k) { dk = qq; break; }
            lval[p] = s / lval[dk];
        }
        if (di < 0) return -(i + 1);
        double d = val[di];
        for (int p = si; p < di; ++p) d -= lval[p] * lval[p];
        if (!(d > 0.0)) return -(i + 1);
        lval[di] = std::sqrt(d);
    }
    return 0;
}

// Zero-fill ILU (IKJ form), the factorization of cusparseDcsrilu02 (preconditioners.cu:48-134):
// L unit-lower and U upper share the CSR of A.
int orc_ilu0(int n, const int* rowptr, const int* col, const double* val, double* lu)
{
    std::memcpy(lu, val, sizeof(double) * rowptr[n]);
    for (int i = 0; i < n; ++i) {
        const int si = rowptr[i], ei = rowptr[i + 1];
        for (int p = si; p < ei; ++p) {
            const int k = col[p];
            if (k >= i) break;
            const int dk = find_col(rowptr, col, k, k);
            lu[p] = lu[p] / lu[dk];
            const double lik = lu[p];
            int q = dk + 1;
            const int ek = rowptr[k + 1];
            for (int pp = p + 1; pp < ei; ++pp) {
                const int j = col[pp];
                while (q < ek && col[q] < j) ++q;
                if (q < ek && col[q] == j) lu[pp] -= lik * lu[q];
            }
        }
        const int di = find_col(rowptr, col, i, i);
        if (di < 0 || lu[di] == 0.0) return -(i + 1);
    }
    return 0;
}

// L y = r, then L^T z = y (SpSV on L and on its transpose, preconditioners.cu:203-212).
void orc_ic0_apply(int n, const int* rowptr, const int* col, const double* l, const double* r, double* z)
{
    std::vector<double> y(n);
    for (int i = 0; i < n; ++i) {
        double s = r[i];
        int di = -1;
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            const int k = col[p];
            if (k < i) s -= l[p] * y[k];
            else if (k == i) { di = p; break; }
        }
        y[i] = s / l[di];
    }
    // column access of L through the symmetric pattern: entry (i,k), k > i, mirrors L_ki at (k,i)
    for (int i = n - 1; i >= 0; --i) {
        double s = y[i];
        int di = -1;
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            const int k = col[p];
            if (k == i) di = p;
            else if (k > i) s -= l[find_col(rowptr, col, k, i)] * z[k];
        }
        z[i] = s / l[di];
    }
}

// L (unit) y = r, then U z = y (preconditioners.cu:128-133).
void orc_ilu0_apply(int n, const int* rowptr, const int* col, const double* lu, const double* r, double* z)
{
    std::vector<double> y(n);
    for (int i = 0; i < n; ++i) {
        double s = r[i];
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            if (col[p] >= i) break;
            s -= lu[p] * y[col[p]];
        }
        y[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = y[i];
        int di = -1;
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            const int k = col[p];
            if (k == i) di = p;
            else if (k > i) s -= lu[p] * z[k];
        }
        z[i] = s / lu[di];
    }
}

// SolverCG::solve (src/linear_algebra/linear_solver.cu:301-376): warm start from x, preconditioner
// re-initialised per solve, stop when (r, M^-1 r) < tol^2 (tol_mode 0, reference) or
// (r, M^-1 r) < tol^2 (b, M^-1 b) (tol_mode 1; absolute when b = 0). x holds the last iterate.
int orc_pcg(int n, const int* rowptr, const int* col, const double* val, double* x, const double* b,
            double tol, int maxit, int tol_mode, int precond, orc_solve_stats* st)
{
    Precond M;
    M.setup(precond, n, rowptr, col, val);
    std::vector<double> r(n), z(n), p(n), Ap(n);
    orc_spmv(n, rowptr, col, val, x, Ap.data());
    for (int i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
    M.apply(z.data(), r.data());
    p = z;
    double gamma = dot(n, r.data(), z.data());
    double thr = tol * tol;
    if (tol_mode == 1) {
        std::vector<double> mb(n);
        M.apply(mb.data(), b);
        const double ref = dot(n, b, mb.data());
        if (ref > 0.0) thr *= ref;   // a zero right-hand side falls back to the absolute test
    }
    st->converged = 0; st->iterations = 0; st->nan_detected = 0;
    if (gamma < thr) { st->converged = 1; st->final_residual = std::sqrt(gamma); return 1; }
    int it = 0;
    while (it < maxit) {
        ++it;
        orc_spmv(n, rowptr, col, val, p.data(), Ap.data());
        const double pAp = dot(n, p.data(), Ap.data());
        const double alpha = gamma / pAp;
#pragma omp parallel for schedule(static) if (n > 8192)
        for (int i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
        M.apply(z.data(), r.data());
        const double gnew = dot(n, r.data(), z.data());
        if (std::isnan(gnew)) { st->nan_detected = 1; gamma = gnew; break; }
        if (gnew < thr) { gamma = gnew; st->converged = 1; break; }
        const double beta = gnew / gamma;
#pragma omp parallel for schedule(static) if (n > 8192)
        for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        gamma = gnew;
    }
    st->iterations = it;
    st->final_residual = std::sqrt(gamma);
    return st->converged;
}

// SolverGMRES::solve (linear_solver.cu:421-521): left preconditioning, modified Gram-Schmidt,
// Givens rotations (updateHcolumn :40-66), stop when |beta_{k+1}| < tol (tol_mode 0) or
// < tol * ||M^-1 b|| (tol_mode 1). restart >= maxit reproduces the reference (no restart, x untouched
// unless converged), and so does restart <= 0 (the library's code for it); restart < maxit is
// GMRES(m): x is updated at the end of every cycle.
// GivensRotation uses 1/sqrt (the reference's rsqrt rounding is not reproducible).
int orc_gmres(int n, const int* rowptr, const int* col, const double* val, double* x, const double* b,
              double tol, int maxit, int restart, int tol_mode, int precond, orc_solve_stats* st)
{
    if (maxit > n) maxit = n;
    if (restart <= 0) restart = maxit;
    const int m = restart < maxit ? restart : maxit;
    const bool restarted = restart < maxit;
    Precond M;
    M.setup(precond, n, rowptr, col, val);
    std::vector<double> V((size_t)n * (m + 1)), H((size_t)m * m, 0.0), cs(m), sn(m), beta(m + 1), w(n);
    double thr = tol;
    if (tol_mode == 1) {
        M.apply(w.data(), b);
        const double ref = std::sqrt(dot(n, w.data(), w.data()));
        if (ref > 0.0) thr = tol * ref;   // a zero right-hand side falls back to the absolute test
    }
    st->converged = 0; st->iterations = 0; st->nan_detected = 0; st->final_residual = 0.0;
    int it = 0;
    while (true) {
        double* v0 = V.data();
        orc_spmv(n, rowptr, col, val, x, v0);
        for (int i = 0; i < n; ++i) v0[i] = b[i] - v0[i];
        M.apply(v0, v0);
        std::fill(beta.begin(), beta.end(), 0.0);
        beta[0] = std::sqrt(dot(n, v0, v0));
        st->final_residual = beta[0];
        if (beta[0] < thr) { st->converged = 1; break; }
        for (int i = 0; i < n; ++i) v0[i] = v0[i] / beta[0];
        int k = 0;
        bool conv = false;
        for (; k < m && it < maxit; ++k) {
            ++it;
            double* vk = V.data() + (size_t)n * k;
            double* vk1 = vk + n;
            orc_spmv(n, rowptr, col, val, vk, vk1);
            M.apply(vk1, vk1);
            double* h = H.data() + (size_t)m * k;
            for (int i = 0; i <= k; ++i) {
                const double* vi = V.data() + (size_t)n * i;
                h[i] = dot(n, vk1, vi);
#pragma omp parallel for schedule(static) if (n > 8192)
                for (int r = 0; r < n; ++r) vk1[r] -= h[i] * vi[r];
            }
            const double hn = std::sqrt(dot(n, vk1, vk1));
#pragma omp parallel for schedule(static) if (n > 8192)
            for (int r = 0; r < n; ++r) vk1[r] = vk1[r] / hn;
            for (int i = 0; i < k; ++i) {
                const double t = cs[i] * h[i] + sn[i] * h[i + 1];
                h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
                h[i] = t;
            }
            const double coeff = 1.0 / std::sqrt(h[k] * h[k] + hn * hn);
            cs[k] = h[k] * coeff;
            sn[k] = hn * coeff;
            h[k] = cs[k] * h[k] + sn[k] * hn;
            beta[k + 1] = -sn[k] * beta[k];
            beta[k] *= cs[k];
            const double res = std::fabs(beta[k + 1]);
            st->final_residual = res;
            if (std::isnan(res)) { st->nan_detected = 1; break; }
            if (res < thr) { conv = true; ++k; break; }
        }
        if (st->nan_detected) break;
        if (conv || restarted) {
            // back substitution H y = beta (cublasDtrsv, upper, non-unit), then x += V y (cublasDgemv)
            std::vector<double> y(beta.begin(), beta.begin() + k);
            for (int i = k - 1; i >= 0; --i) {
                double s = y[i];
                for (int j = i + 1; j < k; ++j) s -= H[(size_t)m * j + i] * y[j];
                y[i] = s / H[(size_t)m * i + i];
            }
#pragma omp parallel for schedule(static) if (n > 8192)
            for (int r = 0; r < n; ++r) {
                double s = 0.0;
                for (int j = 0; j < k; ++j) s += V[(size_t)n * j + r] * y[j];
                x[r] = s + x[r];
            }
        }
        if (conv) { st->converged = 1; break; }
        if (it >= maxit || !restarted) break;
    }
    st->iterations = it;
    return st->converged;
}

// PCG with zero-fill incomplete Cholesky on a matrix whose rows come in colour blocks
// [cptr[c], cptr[c+1]) with no entries between two rows of one colour (a multicolour ordering, as the GPU
// library's pressure preconditioner uses): the factorization, the forward and the backward sweep each run
// colour by colour, the rows of one colour in parallel. Each row's arithmetic is that of orc_ic0 and
// orc_ic0_apply. The CPU baseline's pressure solve.
int orc_pcg_coloured(int n, const int* rowptr, const int* col, const double* val, double* x, const double* b,
                     double tol, int maxit, int tol_mode, int ncol, const int* cptr, orc_solve_stats* st)
{
    const int nnz = rowptr[n];
    std::vector<double> l(nnz, 0.0);
    std::vector<int> diag(n), mirror(nnz, -1);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        diag[i] = find_col(rowptr, col, i, i);
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p)
            if (col[p] > i) mirror[p] = find_col(rowptr, col, col[p], i);
    }
    int bad = 0;
    for (int c = 0; c < ncol; ++c) {
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : bad)
        for (int i = cptr[c]; i < cptr[c + 1]; ++i) {
            const int si = rowptr[i], di = diag[i];
            if (di < 0) { ++bad; continue; }
            for (int p = si; p < di; ++p) {
                const int k = col[p];
                double s = val[p];
                int q = rowptr[k];
                const int pk = rowptr[k + 1];
                for (int pp = si; pp < p; ++pp) {
                    const int j = col[pp];
                    while (q < pk && col[q] < j) ++q;
                    if (q < pk && col[q] == j) s -= l[pp] * l[q];
                }
                l[p] = s / l[diag[k]];
            }
            double d = val[di];
            for (int p = si; p < di; ++p) d -= l[p] * l[p];
            if (!(d > 0.0)) { ++bad; continue; }
            l[di] = std::sqrt(d);
        }
    }
    st->converged = 0; st->iterations = 0; st->nan_detected = 0; st->final_residual = 0.0;
    if (bad) return 0;
    std::vector<double> y(n);
    auto apply = [&](double* z, const double* r) {
        for (int c = 0; c < ncol; ++c) {
#pragma omp parallel for schedule(static)
            for (int i = cptr[c]; i < cptr[c + 1]; ++i) {
                double s = r[i];
                for (int p = rowptr[i]; p < diag[i]; ++p) s -= l[p] * y[col[p]];
                y[i] = s / l[diag[i]];
            }
        }
        for (int c = ncol - 1; c >= 0; --c) {
#pragma omp parallel for schedule(static)
            for (int i = cptr[c]; i < cptr[c + 1]; ++i) {
                double s = y[i];
                for (int p = diag[i] + 1; p < rowptr[i + 1]; ++p) s -= l[mirror[p]] * z[col[p]];
                z[i] = s / l[diag[i]];
            }
        }
    };
    std::vector<double> r(n), z(n), p(n), Ap(n);
    orc_spmv(n, rowptr, col, val, x, Ap.data());
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
    apply(z.data(), r.data());
    p = z;
    double gamma = dot(n, r.data(), z.data());
    double thr = tol * tol;
    if (tol_mode == 1) {
        std::vector<double> mb(n);
        apply(mb.data(), b);
        const double ref = dot(n, b, mb.data());
        if (ref > 0.0) thr *= ref;
    }
    if (gamma < thr) { st->converged = 1; st->final_residual = std::sqrt(gamma); return 1; }
    int it = 0;
    while (it < maxit) {
        ++it;
        orc_spmv(n, rowptr, col, val, p.data(), Ap.data());
        const double alpha = gamma / dot(n, p.data(), Ap.data());
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
        apply(z.data(), r.data());
        const double gnew = dot(n, r.data(), z.data());
        if (std::isnan(gnew)) { st->nan_detected = 1; gamma = gnew; break; }
        if (gnew < thr) { gamma = gnew; st->converged = 1; break; }
        const double beta = gnew / gamma;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        gamma = gnew;
    }
    st->iterations = it;
    st->final_residual = std::sqrt(gamma);
    return st->converged;
}

}  // extern "C"
