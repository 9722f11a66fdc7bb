// Oracle: ParticleHandler2D restated sequentially (TEST INFRASTRUCTURE).
// Follows src/particles/particle_handler_2d.cu and src/particles/particle_2d.cu.
//
// Ordering contract (the reference order depends on device atomics and is not reproducible):
//  * seeding writes particle k of cell c at index c*ppc + k with ID c*ppc + k;
//  * a particle found outside its cell's 1-ring is removed by stable compaction;
//  * refill particles are appended in (cell, subcell) order with consecutive new IDs;
//  * after the census/refill the array is stably sorted by cell index.
// Projection groups the reference's order-free sum as: per cell, particles in array order; per node,
// incident cells in ascending order.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

struct orc_particles {
    int nv = 0, nt = 0;
    std::vector<double> xy, invJ;
    std::vector<int> tri, nbr_off, nbr_idx;
    std::vector<int> inc_off, inc_cell, inc_loc;   // node -> (cell, local index), cells ascending
    int level = 2;
    double centers[16][3];
    double step = 0.5;
    std::vector<int> cell;
    std::vector<double> lam, pos, vel;
    std::vector<uint32_t> id;
    uint32_t next_id = 0;
    long deleted_last = 0, added_last = 0;
};

extern "C" {

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11).
void orc_philox4x32(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// ParticleHandler2D constructor (particle_handler_2d.cu:238-274): L^2 sub-triangle centroids in
// barycentric space, row i from the top holding 2i+1 triangles; center = ONE_THIRD * (v0 + v1 + v2).
void orc_subcell_centers(int level, double* out)
{
    const int L = std::max(std::min(level, 4), 1);
    const double dx = 1.0 / L;
    int k = -1;
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < 2 * i + 1; ++j) {
            const double xmin = (j / 2) * dx, xmax = xmin + dx;
            const double ymin = (L - 1 - i) * dx, ymax = ymin + dx;
            double v[3][3];
            v[0][0] = xmin; v[0][1] = ymax; v[0][2] = 1.0 - xmin - ymax;
            v[1][0] = (j % 2 == 0) ? xmin : xmax;
            v[1][1] = (j % 2 == 0) ? ymin : ymax;
            v[1][2] = 1.0 - v[1][0] - v[1][1];
            v[2][0] = xmax; v[2][1] = ymin; v[2][2] = 1.0 - xmax - ymin;
            ++k;
            for (int d = 0; d < 3; ++d) out[3 * k + d] = (v[0][d] + v[1][d] + v[2][d]) * 0.3333333333333333;
        }
}

// determineSubcell (particle_handler_2d.cu:10-33). Deviation (recorded): the row/column indices are
// clamped to the triangle (i <= L-1, j <= i); the reference indexes past the cell's counters when a
// particle sits within the 2e-6 tolerance outside the bottom or right edge.
int orc_determine_subcell(int L, double l1, double l2, double l3)
{
    const double step = 1.0 / L;
    int i = (int)((1.0 - l2) * L);
    if (i > L - 1) i = L - 1;
    if (i < 0) i = 0;
    int res = i * i;
    int j = (int)(l1 * L);
    if (j > i) j = i;
    if (j < 0) j = 0;
    res += 2 * j;
    if (j != i) {
        const double x = step * (j + 1);
        const double y = 1.0 - step * (i + 1);
        const double z = 1.0 - x - y;
        if (l3 < z) ++res;
    }
    return res;
}

orc_particles* orc_ph_create(int nv, const double* xy, int nt, const int* tri, const double* invJ,
                             const int* nbr_off, const int* nbr_idx, int level)
{
    orc_particles* ph = new orc_particles;
    ph->nv = nv; ph->nt = nt;
    ph->xy.assign(xy, xy + 2 * (size_t)nv);
    ph->tri.assign(tri, tri + 3 * (size_t)nt);
    ph->invJ.assign(invJ, invJ + 4 * (size_t)nt);
    ph->nbr_off.assign(nbr_off, nbr_off + nt + 1);
    ph->nbr_idx.assign(nbr_idx, nbr_idx + nbr_off[nt]);
    ph->level = std::max(std::min(level, 4), 1);
    ph->step = 1.0 / ph->level;
    orc_subcell_centers(ph->level, &ph->centers[0][0]);
    ph->inc_off.assign(nv + 1, 0);
    for (int c = 0; c < nt; ++c)
        for (int i = 0; i < 3; ++i) ph->inc_off[tri[3 * c + i] + 1]++;
    for (int v = 0; v < nv; ++v) ph->inc_off[v + 1] += ph->inc_off[v];
    ph->inc_cell.resize(3 * (size_t)nt);
    ph->inc_loc.resize(3 * (size_t)nt);
    std::vector<int> fill(ph->inc_off.begin(), ph->inc_off.end() - 1);
    for (int c = 0; c < nt; ++c)
        for (int i = 0; i < 3; ++i) {
            const int v = tri[3 * c + i];
            ph->inc_cell[fill[v]] = c;
            ph->inc_loc[fill[v]] = i;
            fill[v]++;
        }
    return ph;
}

void orc_ph_destroy(orc_particles* ph) { delete ph; }

static inline void local_to_global(const orc_particles* ph, int c, const double* L, double* out)
{
    const int* t = &ph->tri[3 * (size_t)c];
    const double* v0 = &ph->xy[2 * (size_t)t[0]];
    const double* v1 = &ph->xy[2 * (size_t)t[1]];
    const double* v2 = &ph->xy[2 * (size_t)t[2]];
    out[0] = L[0] * v0[0] + L[1] * v1[0] + L[2] * v2[0];
    out[1] = L[0] * v0[1] + L[1] * v1[1] + L[2] * v2[1];
}

// uniform double in [0,1) from two 32-bit words (53 random bits)
static inline double u01(uint32_t a, uint32_t b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// seedParticles / kSeedParticlesIntoCell (particle_handler_2d.cu:35-52, :304-320).
// mode 0: particle k at the centroid of subcell k (ppc = L^2, the reference). mode 1: ppc particles at
// barycentric coordinates (1-sqrt(r1), sqrt(r1)(1-r2), sqrt(r1) r2), (r1, r2) from Philox4x32-10 with
// key (seed lo, seed hi) and counter (cell, k, 0, 0) -- uniform over the triangle.
void orc_ph_seed(orc_particles* ph, int mode, int ppc, uint64_t seed)
{
    if (mode == 0) ppc = ph->level * ph->level;
    const size_t P = (size_t)ph->nt * ppc;
    ph->cell.resize(P); ph->lam.resize(3 * P); ph->pos.resize(2 * P); ph->vel.assign(2 * P, 0.0); ph->id.resize(P);
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int c = 0; c < ph->nt; ++c)
        for (int k = 0; k < ppc; ++k) {
            const size_t p = (size_t)c * ppc + k;
            double L[3];
            if (mode == 0) {
                L[0] = ph->centers[k][0]; L[1] = ph->centers[k][1]; L[2] = ph->centers[k][2];
            } else {
                const uint32_t ctr[4] = {(uint32_t)c, (uint32_t)k, 0u, 0u};
                uint32_t r[4];
                orc_philox4x32(ctr, key, r);
                const double s = std::sqrt(u01(r[0], r[1]));
                const double r2 = u01(r[2], r[3]);
                L[0] = 1.0 - s; L[1] = s * (1.0 - r2); L[2] = s * r2;
            }
            ph->cell[p] = c;
            for (int d = 0; d < 3; ++d) ph->lam[3 * p + d] = L[d];
            local_to_global(ph, c, L, &ph->pos[2 * p]);
            ph->id[p] = (uint32_t)p;
        }
    ph->next_id = (uint32_t)P;
}

// initParticleVelocity -> kCorrectParticleVelocity with velocityOld == nullptr (particle_handler_2d.cu:72-88)
void orc_ph_init_velocity(orc_particles* ph, const double* u0, const double* u1)
{
    const size_t P = ph->cell.size();
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < P; ++p) {
        const int* t = &ph->tri[3 * (size_t)ph->cell[p]];
        double ix = 0.0, iy = 0.0;
        for (int i = 0; i < 3; ++i) {
            ix += ph->lam[3 * p + i] * (u0[t[i]] - 0.0);
            iy += ph->lam[3 * p + i] * (u1[t[i]] - 0.0);
        }
        ph->vel[2 * p] = ph->vel[2 * p] + ix;
        ph->vel[2 * p + 1] = ph->vel[2 * p + 1] + iy;
    }
}

// Particle2D::isInsideCell (particle_2d.cu:14-22) with GEOMETRY::transformGlobalToLocal and
// isPointInsideUnitTriangle (geometry.cuh:14-46): inclusive tolerance 2e-6.
static inline bool inside(const orc_particles* ph, int c, const double* x, double* L)
{
    const double* J = &ph->invJ[4 * (size_t)c];
    const int v3 = ph->tri[3 * (size_t)c + 2];
    const double dx = x[0] - ph->xy[2 * (size_t)v3], dy = x[1] - ph->xy[2 * (size_t)v3 + 1];
    L[0] = J[0] * dx + J[2] * dy;
    L[1] = J[1] * dx + J[3] * dy;
    L[2] = 1.0 - L[0] - L[1];
    for (int d = 0; d < 3; ++d)
        if (L[d] > 1.0 + 2e-6 || L[d] < -2e-6) return false;
    return true;
}

static void compact(orc_particles* ph, const std::vector<char>& keep)
{
    size_t w = 0;
    const size_t P = ph->cell.size();
    for (size_t p = 0; p < P; ++p) {
        if (!keep[p]) continue;
        if (w != p) {
            ph->cell[w] = ph->cell[p];
            for (int d = 0; d < 3; ++d) ph->lam[3 * w + d] = ph->lam[3 * p + d];
            for (int d = 0; d < 2; ++d) { ph->pos[2 * w + d] = ph->pos[2 * p + d]; ph->vel[2 * w + d] = ph->vel[2 * p + d]; }
            ph->id[w] = ph->id[p];
        }
        ++w;
    }
    ph->cell.resize(w); ph->lam.resize(3 * w); ph->pos.resize(2 * w); ph->vel.resize(2 * w); ph->id.resize(w);
}

// advectParticles (particle_handler_2d.cu:328-342): per substep kAdvectParticles (:54-70) with the lambda
// of the last relocation, then relocation sortParticlesInCells (:363-395: own cell, then the vertex
// neighbours in ascending order, else delete); then checkParticleDistribution (:397-423: every empty
// subcell of every cell gets a particle at its centroid with grid-interpolated velocity); then the
// particle array is stably ordered by cell.
void orc_ph_advect(orc_particles* ph, const double* u0, const double* u1, double dt, int substeps)
{
    const double dts = dt / substeps;
    ph->deleted_last = 0;
    for (int s = 0; s < substeps; ++s) {
        const size_t P = ph->cell.size();
#pragma omp parallel for schedule(static)
        for (size_t p = 0; p < P; ++p) {
            const int* t = &ph->tri[3 * (size_t)ph->cell[p]];
            double ax = 0.0, ay = 0.0;
            for (int i = 0; i < 3; ++i) {
                ax += ph->lam[3 * p + i] * u0[t[i]];
                ay += ph->lam[3 * p + i] * u1[t[i]];
            }
            ph->pos[2 * p] = ph->pos[2 * p] + dts * ax;
            ph->pos[2 * p + 1] = ph->pos[2 * p + 1] + dts * ay;
        }
        std::vector<char> keep(P, 1);
        long deleted = 0;
#pragma omp parallel for schedule(static) reduction(+ : deleted)
        for (size_t p = 0; p < P; ++p) {
            double L[3];
            const int c = ph->cell[p];
            if (inside(ph, c, &ph->pos[2 * p], L)) {
                for (int d = 0; d < 3; ++d) ph->lam[3 * p + d] = L[d];
                continue;
            }
            bool found = false;
            for (int k = ph->nbr_off[c]; k < ph->nbr_off[c + 1]; ++k) {
                const int nb = ph->nbr_idx[k];
                if (inside(ph, nb, &ph->pos[2 * p], L)) {
                    for (int d = 0; d < 3; ++d) ph->lam[3 * p + d] = L[d];
                    ph->cell[p] = nb;
                    found = true;
                    break;
                }
            }
            if (!found) { keep[p] = 0; ++deleted; }
        }
        ph->deleted_last += deleted;
        compact(ph, keep);
    }
    // census and refill
    const int L = ph->level, ns = L * L;
    std::vector<int> count((size_t)ph->nt * ns, 0);
    const size_t P = ph->cell.size();
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < P; ++p) {
        const size_t k = (size_t)ph->cell[p] * ns + orc_determine_subcell(L, ph->lam[3 * p], ph->lam[3 * p + 1], ph->lam[3 * p + 2]);
#pragma omp atomic
        count[k]++;
    }
    ph->added_last = 0;
    for (int c = 0; c < ph->nt; ++c) {
        const int* t = &ph->tri[3 * (size_t)c];
        for (int i = 0; i < ns; ++i) {
            if (count[(size_t)c * ns + i] != 0) continue;
            const double* Lc = ph->centers[i];
            double x[2];
            local_to_global(ph, c, Lc, x);
            double vx = 0.0, vy = 0.0;
            for (int j = 0; j < 3; ++j) { vx += Lc[j] * u0[t[j]]; vy += Lc[j] * u1[t[j]]; }
            ph->cell.push_back(c);
            for (int d = 0; d < 3; ++d) ph->lam.push_back(Lc[d]);
            ph->pos.push_back(x[0]); ph->pos.push_back(x[1]);
            ph->vel.push_back(vx); ph->vel.push_back(vy);
            ph->id.push_back(ph->next_id++);
            ph->added_last++;
        }
    }
    // stable order by cell (a counting sort)
    const size_t Q = ph->cell.size();
    std::vector<size_t> perm(Q), start((size_t)ph->nt + 1, 0);
    for (size_t p = 0; p < Q; ++p) start[(size_t)ph->cell[p] + 1]++;
    for (int c = 0; c < ph->nt; ++c) start[c + 1] += start[c];
    for (size_t p = 0; p < Q; ++p) perm[start[ph->cell[p]]++] = p;
    std::vector<int> c2(Q); std::vector<double> l2(3 * Q), x2(2 * Q), v2(2 * Q); std::vector<uint32_t> i2(Q);
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < Q; ++k) {
        const size_t p = perm[k];
        c2[k] = ph->cell[p]; i2[k] = ph->id[p];
        for (int d = 0; d < 3; ++d) l2[3 * k + d] = ph->lam[3 * p + d];
        for (int d = 0; d < 2; ++d) { x2[2 * k + d] = ph->pos[2 * p + d]; v2[2 * k + d] = ph->vel[2 * p + d]; }
    }
    ph->cell.swap(c2); ph->lam.swap(l2); ph->pos.swap(x2); ph->vel.swap(v2); ph->id.swap(i2);
}

// projectVelocityOntoGrid (particle_handler_2d.cu:90-115, :350-361): u_i = sum(lambda_i v) / sum(lambda_i),
// grouped per cell then per node (cells ascending). Requires the cell-ordered array left by advect.
void orc_ph_project(orc_particles* ph, double* u0, double* u1)
{
    const size_t P = ph->cell.size();
    std::vector<double> part(9 * (size_t)ph->nt, 0.0);
#pragma omp parallel for schedule(static)
    for (size_t p = 0; p < P; ++p) {
        double* s = &part[9 * (size_t)ph->cell[p]];
        for (int i = 0; i < 3; ++i) {
            const double l = ph->lam[3 * p + i];
#pragma omp atomic
            s[3 * i + 0] += l * ph->vel[2 * p];
#pragma omp atomic
            s[3 * i + 1] += l * ph->vel[2 * p + 1];
#pragma omp atomic
            s[3 * i + 2] += l;
        }
    }
#pragma omp parallel for schedule(static)
    for (int v = 0; v < ph->nv; ++v) {
        double sx = 0.0, sy = 0.0, w = 0.0;
        for (int k = ph->inc_off[v]; k < ph->inc_off[v + 1]; ++k) {
            const double* s = &part[9 * (size_t)ph->inc_cell[k] + 3 * ph->inc_loc[k]];
            sx += s[0]; sy += s[1]; w += s[2];
        }
        u0[v] = sx / w;
        u1[v] = sy / w;
    }
}

// correctParticleVelocity -> kCorrectParticleVelocity (particle_handler_2d.cu:72-88, :344-348)
void orc_ph_correct(orc_particles* ph, const double* u0, const double* u1, const double* uo0, const double* uo1)
{
    const size_t P = ph->cell.size();
    for (size_t p = 0; p < P; ++p) {
        const int* t = &ph->tri[3 * (size_t)ph->cell[p]];
        double ix = 0.0, iy = 0.0;
        for (int i = 0; i < 3; ++i) {
            ix += ph->lam[3 * p + i] * (u0[t[i]] - uo0[t[i]]);
            iy += ph->lam[3 * p + i] * (u1[t[i]] - uo1[t[i]]);
        }
        ph->vel[2 * p] = ph->vel[2 * p] + ix;
        ph->vel[2 * p + 1] = ph->vel[2 * p + 1] + iy;
    }
}

long orc_ph_count(const orc_particles* ph) { return (long)ph->cell.size(); }

void orc_ph_get(const orc_particles* ph, int* cell, double* lam, double* pos, double* vel, uint32_t* id)
{
    const size_t P = ph->cell.size();
    if (cell) std::memcpy(cell, ph->cell.data(), P * sizeof(int));
    if (lam) std::memcpy(lam, ph->lam.data(), 3 * P * sizeof(double));
    if (pos) std::memcpy(pos, ph->pos.data(), 2 * P * sizeof(double));
    if (vel) std::memcpy(vel, ph->vel.data(), 2 * P * sizeof(double));
    if (id) std::memcpy(id, ph->id.data(), P * sizeof(uint32_t));
}

void orc_ph_set(orc_particles* ph, long count, const int* cell, const double* lam, const double* pos,
                const double* vel, const uint32_t* id, uint32_t next_id)
{
    const size_t P = (size_t)count;
    ph->cell.assign(cell, cell + P);
    ph->lam.assign(lam, lam + 3 * P);
    ph->pos.assign(pos, pos + 2 * P);
    ph->vel.assign(vel, vel + 2 * P);
    ph->id.assign(id, id + P);
    ph->next_id = next_id;
}

void orc_ph_stats(const orc_particles* ph, long* deleted_last, long* added_last)
{
    *deleted_last = ph->deleted_last;
    *added_last = ph->added_last;
}

}  // extern "C"
