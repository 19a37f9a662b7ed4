// engine.hip -- MI355X (gfx950) DEM step engine behind the C-ABI of
// include/xdem.h.
//
// One step of every packing is five launches on the engine's stream:
//   k_bin         wrap positions (modifyParticlePosition), cell index, fill
//                 fixed-capacity cell bins and the two ghost rows the
//                 Lees-Edwards boundary needs (findPossibleContacts binning)
//   k_candidates  3x3 stencil, candidate filters, image offsets, bounding
//                 circle test; block-compacted pair list; carries forward
//                 history entries whose pair is no longer a candidate
//   k_pairs       one lane per pair: ZPM / circle geometry, history lookup
//                 in the previous contact table, contact model
//                 (BaseParticle::calculateContactForceAndTorque)
//   k_gather      per-particle force/torque sums in a fixed order
//                 (collectForceAndTorque)
//   k_integrate   background-velocity removal, integrator, clumps, LE shift
//                 (takeTimeIntegral, updateLeesEdwardBoundary)
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/xdem.h"
#include "dem_math.h"

using namespace xdem;

namespace {

constexpr int kBlock = 256;
constexpr int kMaxCand = 200;      // src/BaseWorld.h:11
constexpr int kMaxPairs = 32;      // bounding-circle pairs j>i per particle
constexpr int kInCap = 32;         // contacts per particle held by partners
constexpr int kImgShift = 30;
constexpr int kImgStale = 3;

struct Pack {
    double xmin, xmax, ymin, ymax, xlen, ylen, lenx, leny, hx, hy;
    double dt, kn, kt, mu, factor, gdamp, rate;
    double shift, shift_vel, strain;
    int nx, ny, bin_base, world, p_begin, p_end, newton_cap, strict;
};

struct Parts {
    double *x, *y, *th, *vx, *vy, *w;
    const double *r, *ar, *m, *im, *iin;
    double *cbx, *cby;
    const int* up;
    const int8_t* kind;
    const int8_t* flags;
    const int* pack;
    double *fx, *fy, *tq;
    int* ncont;
    int *bin_of, *gbin_of;
    int* in_cnt;
    int2* in_ref;
    int n;
};

struct Table {
    int *pr_i, *pr_j, *ct_j;
    double *xix, *xiy, *nx, *ny, *c;
    double *fnx, *fny, *fsx, *fsy, *t1, *t2, *bx, *by, *px, *py;
    int *row_start, *row_cnt;
    int* counter;
    int cap;
};

struct Bins {
    int* count;
    int* list;
    int cap;
    int nbins;
};

struct Counters {
    unsigned long long cand, coarse, newton, rescues, fallbacks;
};

__device__ __forceinline__ void raise_err(uint32_t* err, uint32_t bit) {
    atomicOr(err, bit);
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ void count_add(unsigned long long* dst, unsigned long long v) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, v);
}

// Image offset of P2 for a pair found in bin row `img`:
// 0: minimum image (src/PeriodicBoundaryWorld.cpp:97-101,
//    src/LeesEdwardsBoundaryWorld.cpp:218-222);
// 1/2: Lees-Edwards ghost row above/below (src/LeesEdwardsBoundaryWorld.cpp:233-279).
__device__ __forceinline__ void image_offset(const Pack& K, int img, double xi, double yi,
                                             double xj, double yj, double& ox, double& oy,
                                             double& ovx) {
    ox = 0.0; oy = 0.0; ovx = 0.0;
    if (img == 0) {
        if (xj - xi > K.xlen / 2.0) ox = -K.xlen;
        else if (xj - xi < -K.xlen / 2.0) ox = K.xlen;
        if (yj - yi > K.ylen / 2.0) oy = -K.ylen;
        else if (yj - yi < -K.ylen / 2.0) oy = K.ylen;
    } else {
        ox = img == 1 ? K.shift : -K.shift;
        oy = img == 1 ? K.ylen : -K.ylen;
        double tx = ox + xj;
        if (tx < K.xmin) { ox += K.xlen; tx += K.xlen; }
        else if (tx > K.xmax) { ox += -K.xlen; tx -= K.xlen; }
        if (tx - xi > K.xlen / 2.0) ox += -K.xlen;
        else if (tx - xi < -K.xlen / 2.0) ox += K.xlen;
        ovx = img == 1 ? K.shift_vel : -K.shift_vel;
    }
}

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
k_bin(Parts P, const Pack* __restrict__ packs, Bins B, uint32_t* err) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= P.n || *err) return;
    const Pack K = packs[P.pack[p]];
    double x = P.x[p], y = P.y[p];
    const double x0 = x, y0 = y;
    // modifyParticlePosition: LE src/LeesEdwardsBoundaryWorld.cpp:304-333,
    // periodic src/PeriodicBoundaryWorld.cpp:129-142
    if (K.world == WORLD_LE) {
        if (y < K.ymin) { y += K.ylen; x += K.shift; }
        else if (y > K.ymax) { y += -K.ylen; x += -K.shift; }
    } else {
        if (y < K.ymin) y += K.ylen;
        else if (y > K.ymax) y += -K.ylen;
    }
    if (x < K.xmin) x += K.xlen;
    else if (x > K.xmax) x += -K.xlen;
    if (x != x0 || y != y0) { P.x[p] = x; P.y[p] = y; }
    P.in_cnt[p] = 0;
    // cell, src/PeriodicBoundaryWorld.cpp:32-33
    int ix = (int)floor((x - K.xmin) / K.lenx);
    int iy = (int)floor((y - K.ymin) / K.leny);
    ix = min(max(ix, 0), K.nx - 1);
    iy = min(max(iy, 0), K.ny - 1);
    const int bin = K.bin_base + (iy + 1) * K.nx + ix;
    int rank = atomicAdd(&B.count[bin], 1);
    if (rank < B.cap) B.list[(size_t)bin * B.cap + rank] = p;
    else raise_err(err, ERR_BIN_OVERFLOW);
    P.bin_of[p] = bin;
    int gbin = -1;
    if (iy == 0 || iy == K.ny - 1) {
        // ghost rows: bottom row at x+S feeds the row above the grid, top row
        // at x-S the row below (src/LeesEdwardsBoundaryWorld.cpp:75-88); for
        // the periodic world the unshifted copy reproduces the y wrap of the
        // stencil (src/PeriodicBoundaryWorld.cpp:64-65)
        int gx = ix;
        if (K.world == WORLD_LE) {
            double t = iy == 0 ? x + K.shift : x - K.shift;
            if (t < K.xmin) t += K.xlen;
            else if (t > K.xmax) t -= K.xlen;
            gx = (int)floor((t - K.xmin) / K.lenx);
            gx = min(max(gx, 0), K.nx - 1);
        }
        const int row = iy == 0 ? K.ny + 1 : 0;
        gbin = K.bin_base + row * K.nx + gx;
        rank = atomicAdd(&B.count[gbin], 1);
        if (rank < B.cap) B.list[(size_t)gbin * B.cap + rank] = p;
        else raise_err(err, ERR_BIN_OVERFLOW);
    }
    P.gbin_of[p] = gbin;
}

// Is P2 = j in P1 = i's 3x3 stencil this step (same filters as k_candidates)?
__device__ __forceinline__ bool in_stencil(const Pack& K, int bin_i, int bin_j) {
    const int li = bin_i - K.bin_base, lj = bin_j - K.bin_base;
    const int ri = li / K.nx, ci = li - ri * K.nx;
    const int rj = lj / K.nx, cj = lj - rj * K.nx;
    if (rj < ri - 1 || rj > ri + 1) return false;
    int dc = cj - ci;
    if (dc == K.nx - 1) dc = -1;
    else if (dc == -(K.nx - 1)) dc = 1;
    return dc >= -1 && dc <= 1;
}

__device__ __forceinline__ bool is_candidate(const Parts& P, const Pack& K, int i, int j) {
    if (K.world == WORLD_PERIODIC &&
        ((P.flags[i] & SKIP_CONTACT_DETECTION) || (P.flags[j] & SKIP_CONTACT_DETECTION)))
        return false;
    if (i >= j || P.up[i] == P.up[j]) return false;
    const int bi = P.bin_of[i];
    if (in_stencil(K, bi, P.bin_of[j])) return true;
    const int g = P.gbin_of[j];
    return g >= 0 && in_stencil(K, bi, g);
}

__global__ void __launch_bounds__(kBlock)
k_candidates(Parts P, const Pack* __restrict__ packs, Bins B, Table old_t, Table new_t,
             uint32_t* err, Counters* cnt) {
    __shared__ int s_pairs[kMaxPairs][kBlock];
    __shared__ int s_scan[kBlock];
    __shared__ int s_base;
    const int tid = threadIdx.x;
    const int i = blockIdx.x * kBlock + tid;
    const bool active = i < P.n && *err == 0;
    int npair = 0, nstale = 0;
    unsigned long long ncand_total = 0;
    if (active) {
        const Pack K = packs[P.pack[i]];
        const bool skip_i = K.world == WORLD_PERIODIC && (P.flags[i] & SKIP_CONTACT_DETECTION);
        if (!skip_i) {
            const double xi = P.x[i], yi = P.y[i], ri = P.r[i];
            const int upi = P.up[i], ki = P.kind[i];
            const int li = P.bin_of[i] - K.bin_base;
            const int row_i = li / K.nx, col_i = li - row_i * K.nx;
            int ncand[3] = {0, 0, 0};
            for (int jx0 = -1; jx0 <= 1; ++jx0) {
                int jx = col_i + jx0;
                if (jx == -1) jx = K.nx - 1;
                else if (jx == K.nx) jx = 0;
                for (int jy0 = -1; jy0 <= 1; ++jy0) {
                    const int row = row_i + jy0;
                    int img = 0;
                    if (K.world == WORLD_LE) img = row == 0 ? 2 : (row == K.ny + 1 ? 1 : 0);
                    const int b = K.bin_base + row * K.nx + jx;
                    const int nb = min(B.count[b], B.cap);
                    const int* lst = B.list + (size_t)b * B.cap;
                    for (int t = 0; t < nb; ++t) {
                        const int j = lst[t];
                        if (i >= j) continue;
                        if (K.world == WORLD_PERIODIC && (P.flags[j] & SKIP_CONTACT_DETECTION)) continue;
                        if (upi == P.up[j]) continue;
                        ncand[img]++;
                        if (ki == KIND_CLUMP) continue;          // never in contact
                        const int kj = P.kind[j];
                        if (kj != ki) { raise_err(err, ERR_UNSUPPORTED); continue; }
                        const double xj = P.x[j], yj = P.y[j];
                        double ox, oy, ovx;
                        image_offset(K, img, xi, yi, xj, yj, ox, oy, ovx);
                        bool pass;
                        if (ki == KIND_ELLIPSE) {
                            // takeCoarseDetection, src/BaseParticle.h:131-132
                            const double qx = xi - (xj + ox), qy = yi - (yj + oy);
                            pass = (ri + P.r[j]) > sqrt(qx * qx + qy * qy);
                        } else {
                            // circles: the contact test itself (src/CircularParticle.cpp:73-80)
                            const double bx = (xj + ox) - xi, by = (yj + oy) - yi;
                            pass = !((ri + P.r[j]) - sqrt(bx * bx + by * by) < 0.0);
                        }
                        if (pass) {
                            if (npair < kMaxPairs) s_pairs[npair][tid] = j | (img << kImgShift);
                            ++npair;
                        }
                    }
                }
            }
            if (ncand[0] >= kMaxCand || ncand[1] >= kMaxCand || ncand[2] >= kMaxCand)
                raise_err(err, ERR_CAND_OVERFLOW);
            ncand_total = (unsigned long long)(ncand[0] + ncand[1] + ncand[2]);
        }
        if (npair > kMaxPairs) { raise_err(err, ERR_ROW_OVERFLOW); npair = kMaxPairs; }
        // history entries of i whose pair is no longer a candidate stay
        // untouched in the reference and keep contributing force
        // (src/BaseWorld.cpp:34-46)
        const int os = old_t.row_start[i], oc = old_t.row_cnt[i];
        for (int e = os; e < os + oc; ++e) {
            const int j = old_t.ct_j[e];
            if (j >= 0 && !is_candidate(P, K, i, j)) ++nstale;
        }
    }
    count_add(&cnt->cand, ncand_total);
    count_add(&cnt->coarse, (unsigned long long)npair);
    // block exclusive scan of entries per particle
    const int total = npair + nstale;
    s_scan[tid] = total;
    __syncthreads();
    for (int o = 1; o < kBlock; o <<= 1) {
        const int v = tid >= o ? s_scan[tid - o] : 0;
        __syncthreads();
        s_scan[tid] += v;
        __syncthreads();
    }
    if (tid == kBlock - 1) {
        const int block_total = s_scan[tid];
        int base = block_total ? atomicAdd(new_t.counter, block_total) : 0;
        if ((long long)base + block_total > new_t.cap) {
            raise_err(err, ERR_PAIR_OVERFLOW);
            base = -1;
        }
        s_base = base;
    }
    __syncthreads();
    if (!active) return;
    const int base = s_base;
    if (base < 0) { new_t.row_start[i] = 0; new_t.row_cnt[i] = 0; return; }
    const int s0 = base + s_scan[tid] - total;
    new_t.row_start[i] = s0;
    new_t.row_cnt[i] = total;
    for (int k = 0; k < npair; ++k) {
        new_t.pr_i[s0 + k] = i;
        new_t.pr_j[s0 + k] = s_pairs[k][tid];
    }
    if (nstale) {
        const Pack K = packs[P.pack[i]];
        int k = npair;
        const int os = old_t.row_start[i], oc = old_t.row_cnt[i];
        for (int e = os; e < os + oc; ++e) {
            const int j = old_t.ct_j[e];
            if (j < 0 || is_candidate(P, K, i, j)) continue;
            const int d = s0 + k++;
            new_t.pr_i[d] = i;
            new_t.pr_j[d] = j | (kImgStale << kImgShift);
            new_t.ct_j[d] = j;
            new_t.xix[d] = old_t.xix[e]; new_t.xiy[d] = old_t.xiy[e];
            new_t.nx[d] = old_t.nx[e];   new_t.ny[d] = old_t.ny[e];
            new_t.c[d] = old_t.c[e];
            new_t.fnx[d] = old_t.fnx[e]; new_t.fny[d] = old_t.fny[e];
            new_t.fsx[d] = old_t.fsx[e]; new_t.fsy[d] = old_t.fsy[e];
            new_t.t1[d] = old_t.t1[e];   new_t.t2[d] = old_t.t2[e];
            new_t.bx[d] = old_t.bx[e];   new_t.by[d] = old_t.by[e];
            new_t.px[d] = old_t.px[e];   new_t.py[d] = old_t.py[e];
            const int q = atomicAdd(&P.in_cnt[j], 1);
            if (q < kInCap) P.in_ref[(size_t)j * kInCap + q] = make_int2(i, d);
            else raise_err(err, ERR_INCOMING_OVERFLOW);
        }
    }
}

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
k_pairs(Parts P, const Pack* __restrict__ packs, Table old_t, Table new_t, uint32_t* err,
        Counters* cnt) {
    if (*err) return;
    const int total = min(*new_t.counter, new_t.cap);
    unsigned long long n_newton = 0, n_rescue = 0;
    for (int e = blockIdx.x * kBlock + threadIdx.x; e < total; e += gridDim.x * kBlock) {
        const int pj = new_t.pr_j[e];
        const int img = (int)((unsigned)pj >> kImgShift);
        const int j = pj & ((1 << kImgShift) - 1);
        const int i = new_t.pr_i[e];
        if (img == kImgStale || i < 0 || i >= P.n || j >= P.n) continue;
        const Pack K = packs[P.pack[i]];
        const double xi = P.x[i], yi = P.y[i], xj = P.x[j], yj = P.y[j];
        double ox, oy, ovx;
        image_offset(K, img, xi, yi, xj, yj, ox, oy, ovx);
        const double xj2 = xj + ox, yj2 = yj + oy;
        const double dx = xj2 - xi, dy = yj2 - yi;
        Geom g;
        int hit;
        if (P.kind[i] == KIND_CIRCLE) {
            hit = circle_contact(xi, yi, P.r[i], xj2, yj2, P.r[j], g) ? 1 : 0;
        } else {
            ZpmSetup S;
            hit = zpm_setup(P.r[i], P.ar[i], P.th[i], P.r[j], P.ar[j], P.th[j], dx, dy, S);
            if (hit) {
                const double qinit = sqrt(S.delta + 1.0);
                double q;
                int it = newton_quartic(S.A, S.B, S.C, S.D, S.E, qinit, K.newton_cap, &q);
                if (it < 0) {
                    if (K.strict || !quartic_bisect(S.A, S.B, S.C, S.D, S.E, qinit, &q)) {
                        raise_err(err, ERR_NEWTON_CAP);
                        new_t.ct_j[e] = -1;
                        continue;
                    }
                    ++n_rescue;
                    it = K.newton_cap;
                }
                n_newton += (unsigned long long)it;
                hit = zpm_finish(S, q, xi, yi, xj2, yj2, dx, dy, g);
                if (hit < 0) { raise_err(err, ERR_ZERO_NORMAL); new_t.ct_j[e] = -1; continue; }
            }
        }
        if (!hit) { new_t.ct_j[e] = -1; continue; }
        // relative velocity, src/BaseParticle.cpp:11-25; LE velocities carry
        // the background term vx + rate*y (src/LeesEdwardsBoundaryWorld.cpp:508-518)
        const bool le = K.world == WORLD_LE;
        const double vxi = le ? P.vx[i] + K.rate * yi : P.vx[i];
        const double vxj = le ? P.vx[j] + K.rate * yj : P.vx[j];
        const double wi = P.w[i], wj = P.w[j];
        const double ax = vxi + (-wi * g.b1y), ay = P.vy[i] + wi * g.b1x;
        const double bx = (vxj + ovx) + (-wj * g.b2y), by = (P.vy[j] + 0.0) + wj * g.b2x;
        const double rvx = ax - bx, rvy = ay - by;
        const double sv = rvx * g.nx + rvy * g.ny;
        const double nrvx = g.nx * sv, nrvy = g.ny * sv;
        const double srvx = rvx - nrvx, srvy = rvy - nrvy;
        // history, src/BaseParticle.cpp:36-82
        const int os = old_t.row_start[i], oc = old_t.row_cnt[i];
        int found = -1;
        for (int t = os; t < os + oc; ++t)
            if (old_t.ct_j[t] == j) { found = t; break; }
        double xix, xiy, c;
        if (found < 0) {
            const double vn = rvx * g.nx + rvy * g.ny;
            if (!isnormal(vn)) { xix = srvx * K.dt; xiy = srvy * K.dt; }
            else {
                const double t = g.overlap / vn;
                const double s = t < K.dt ? t : K.dt;
                xix = srvx * s; xiy = srvy * s;
            }
            // damping constant, src/ContactModel_LinearDissipativeFrictional.cpp:73-82
            const bool si = P.flags[i] & SKIP_TIME_INTEGRAL, sj = P.flags[j] & SKIP_TIME_INTEGRAL;
            if (!si && sj) c = sqrt(P.m[i]) * K.factor;
            else if (si && !sj) c = sqrt(P.m[j]) * K.factor;
            else {
                const double m0 = (P.m[i] * P.m[j]) / (P.m[i] + P.m[j]);
                c = sqrt(m0) * K.factor;
            }
        } else {
            const double onx = old_t.nx[found], ony = old_t.ny[found];
            const double ox0 = old_t.xix[found], oy0 = old_t.xiy[found];
            const double dot = onx * g.nx + ony * g.ny;
            const double det = onx * g.ny - ony * g.nx;
            const double ang = atan2(det, dot);
            double sa, ca;
            sincos(ang, &sa, &ca);
            const double rx = ca * ox0 + (-sa) * oy0;
            const double ry = sa * ox0 + ca * oy0;
            xix = rx + srvx * K.dt;
            xiy = ry + srvy * K.dt;
            c = old_t.c[found];
        }
        // contact model, src/ContactModel_LinearDissipativeFrictional.cpp:3-46
        const double fnx = (-g.nx) * K.kn * g.overlap + (-nrvx) * c;
        const double fny = (-g.ny) * K.kn * g.overlap + (-nrvy) * c;
        const double fmax = sqrt(fnx * fnx + fny * fny) * K.mu;
        const double ex = (-xix) * K.kt, ey = (-xiy) * K.kt;
        double fsx = ex + (-srvx) * c, fsy = ey + (-srvy) * c;
        if (sqrt(fsx * fsx + fsy * fsy) > fmax) {
            const double ls = fsx * fsx + fsy * fsy;
            double ux = 0.0, uy = 0.0;
            if (ls != 0.0) { const double s = sqrt(ls); ux = fsx / s; uy = fsy / s; }
            fsx = ux * fmax; fsy = uy * fmax;
            if (sqrt(ex * ex + ey * ey) > fmax) { xix = (-fsx) / K.kt; xiy = (-fsy) / K.kt; }
            else { xix = (-ex) / K.kt; xiy = (-ey) / K.kt; }
        }
        // src/ContactInformation.h:62-68
        const double fx = fnx + fsx, fy = fny + fsy;
        new_t.ct_j[e] = j;
        new_t.xix[e] = xix; new_t.xiy[e] = xiy;
        new_t.nx[e] = g.nx; new_t.ny[e] = g.ny;
        new_t.c[e] = c;
        new_t.fnx[e] = fnx; new_t.fny[e] = fny;
        new_t.fsx[e] = fsx; new_t.fsy[e] = fsy;
        new_t.t1[e] = g.b1x * fy - g.b1y * fx;
        new_t.t2[e] = g.b2x * (-fy) - g.b2y * (-fx);
        new_t.bx[e] = xi - (xj + ox);
        new_t.by[e] = yi - (yj + oy);
        new_t.px[e] = g.px; new_t.py[e] = g.py;
        const int q = atomicAdd(&P.in_cnt[j], 1);
        if (q < kInCap) P.in_ref[(size_t)j * kInCap + q] = make_int2(i, e);
        else raise_err(err, ERR_INCOMING_OVERFLOW);
    }
    count_add(&cnt->newton, n_newton);
    count_add(&cnt->rescues, n_rescue);
}

// ---------------------------------------------------------------------------
// collectForceAndTorque (src/BaseWorld.cpp:17-57) in the order of a
// single-threaded reference sweep with ordered maps: contacts held by lower
// IDs (ascending), then the particle's own contacts (partner ascending).
__global__ void __launch_bounds__(kBlock)
k_gather(Parts P, Table t, uint32_t* err) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= P.n || *err) return;
    double fx = 0.0, fy = 0.0, tq = 0.0;
    int nc = 0;
    const int nin = min(P.in_cnt[p], kInCap);
    const int2* refs = P.in_ref + (size_t)p * kInCap;
    int last = -1;
    for (int k = 0; k < nin; ++k) {
        int best = 0x7fffffff, be = -1;
        for (int u = 0; u < nin; ++u) {
            const int2 r = refs[u];
            if (r.x > last && r.x < best) { best = r.x; be = r.y; }
        }
        if (be < 0) break;
        last = best;
        fx += -(t.fnx[be] + t.fsx[be]);
        fy += -(t.fny[be] + t.fsy[be]);
        tq += t.t2[be];
        ++nc;
    }
    const int s = t.row_start[p], cnt = t.row_cnt[p];
    last = -1;
    for (int k = 0; k < cnt; ++k) {
        int best = 0x7fffffff, be = -1;
        for (int e = s; e < s + cnt; ++e) {
            const int j = t.ct_j[e];
            if (j > last && j < best) { best = j; be = e; }
        }
        if (be < 0) break;
        last = best;
        fx += t.fnx[be] + t.fsx[be];
        fy += t.fny[be] + t.fsy[be];
        tq += t.t1[be];
        ++nc;
    }
    P.fx[p] = fx; P.fy[p] = fy; P.tq[p] = tq; P.ncont[p] = nc;
}

// takeTimeIntegral (src/BaseWorld.cpp:59-77, src/BaseParticle.cpp:107-118,
// clumps src/GluedParticle.cpp:91-125) after removeBackgroudVelocity
// (src/LeesEdwardsBoundaryWorld.cpp:520-530); then the per-packing LE shift
// update (src/LeesEdwardsBoundaryWorld.cpp:285-296) and clearing of the
// per-step bins and the next table's counter.
__global__ void __launch_bounds__(kBlock)
k_integrate(Parts P, Pack* packs, Bins B, int* next_counter, uint32_t* err) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    for (int b = p; b < B.nbins; b += gridDim.x * kBlock) B.count[b] = 0;
    if (p == 0) *next_counter = 0;
    if (p >= P.n || *err) return;
    const int pk = P.pack[p];
    const Pack K = packs[pk];
    const double g = K.gdamp, dt = K.dt;
    if (K.world == WORLD_LE) {
        const double yb = P.y[p];
        const double vxb = P.vx[p] + K.rate * yb;
        P.vx[p] = vxb - K.rate * yb;
    }
    if (!(P.flags[p] & SKIP_TIME_INTEGRAL)) {
        if (P.kind[p] != KIND_CLUMP) {
            const double im = P.im[p];
            double vx = P.vx[p], vy = P.vy[p];
            vx = (vx * (1.0 - g * dt) + (P.fx[p] * im) * dt) / (1.0 + g * dt);
            vy = (vy * (1.0 - g * dt) + (P.fy[p] * im) * dt) / (1.0 + g * dt);
            P.vx[p] = vx; P.vy[p] = vy;
            P.x[p] += vx * dt;
            P.y[p] += vy * dt;
            double w = P.w[p];
            w = (w * (1.0 - g * dt) + P.tq[p] * P.iin[p] * dt) / (1.0 + g * dt);
            P.w[p] = w;
            P.th[p] += w * dt;
        } else {
            double tfx = 0.0, tfy = 0.0, tt = 0.0;
            int s = p + 1;
            for (; s < P.n && P.up[s] == p; ++s) {
                tfx += P.fx[s]; tfy += P.fy[s];
                tt += P.tq[s];
                tt += P.cbx[s] * P.fy[s] - P.cby[s] * P.fx[s];
            }
            // subs belong to this packing; their LE round trip is done here so
            // that the clump's writes below are the last ones
            if (K.world == WORLD_LE) {
                for (int k = p + 1; k < s; ++k) {
                    const double vxb = P.vx[k] + K.rate * P.y[k];
                    P.vx[k] = vxb - K.rate * P.y[k];
                }
            }
            const double im = P.im[p];
            double vx = P.vx[p], vy = P.vy[p];
            vx = (vx * (1.0 - g * dt) + tfx * im * dt) / (1.0 + g * dt);
            vy = (vy * (1.0 - g * dt) + tfy * im * dt) / (1.0 + g * dt);
            P.vx[p] = vx; P.vy[p] = vy;
            const double x = P.x[p] + vx * dt, y = P.y[p] + vy * dt;
            P.x[p] = x; P.y[p] = y;
            double w = P.w[p];
            w = (w * (1.0 - g * dt) + tt * P.iin[p] * dt) / (1.0 + g * dt);
            P.w[p] = w;
            P.th[p] += w * dt;
            const double ang = w * dt;
            double sa, ca;
            sincos(ang, &sa, &ca);
            for (int k = p + 1; k < s; ++k) {
                const double bx = ca * P.cbx[k] + (-sa) * P.cby[k];
                const double by = sa * P.cbx[k] + ca * P.cby[k];
                P.cbx[k] = bx; P.cby[k] = by;
                P.vx[k] = vx + (-w * by);
                P.vy[k] = vy + w * bx;
                P.x[k] = x + bx;
                P.y[k] = y + by;
                P.w[k] = w;
                P.th[k] += ang;
            }
        }
    }
    if (p == K.p_begin && K.world == WORLD_LE) {
        Pack& W = packs[pk];
        W.shift_vel = K.rate * K.ylen;
        W.strain = K.strain + K.rate * K.dt;
        double sh = K.shift + W.shift_vel * K.dt;
        if (sh > K.xmax) sh -= K.xlen;
        else if (sh < K.xmin) sh += K.xlen;
        W.shift = sh;
    }
}

// Per-particle stress contribution (BaseParticle::getStressFromContactList,
// src/BaseParticle.cpp:143-151; clumps src/GluedParticle.cpp:127-139), own
// contacts in partner order; the host sums particles in index order.
__global__ void __launch_bounds__(kBlock)
k_stress(Parts P, Table t, double* out) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= P.n) return;
    double sxx = 0, sxy = 0, syx = 0, syy = 0;
    if (!(P.flags[p] & SKIP_MACRO_SUM)) {
        const bool clump = P.kind[p] == KIND_CLUMP;
        const int s_end = clump ? P.n : p + 1;
        for (int s = clump ? p + 1 : p; s < s_end; ++s) {
            if (clump && P.up[s] != p) break;
            const double ox = clump ? P.cbx[s] : 0.0, oy = clump ? P.cby[s] : 0.0;
            const int rs = t.row_start[s], cnt = t.row_cnt[s];
            int last = -1;
            for (int k = 0; k < cnt; ++k) {
                int best = 0x7fffffff, be = -1;
                for (int e = rs; e < rs + cnt; ++e) {
                    const int j = t.ct_j[e];
                    if (j > last && j < best) { best = j; be = e; }
                }
                if (be < 0) break;
                last = best;
                const double fx = t.fnx[be] + t.fsx[be], fy = t.fny[be] + t.fsy[be];
                const double bx = clump ? t.bx[be] + ox : t.bx[be];
                const double by = clump ? t.by[be] + oy : t.by[be];
                sxx += fx * bx; sxy += fx * by; syx += fy * bx; syy += fy * by;
            }
        }
    }
    out[4 * (size_t)p + 0] = sxx; out[4 * (size_t)p + 1] = sxy;
    out[4 * (size_t)p + 2] = syx; out[4 * (size_t)p + 3] = syy;
}

// history restart: one contact per thread into the old table's rows
__global__ void k_load_contacts(Table t, int nc, const int* ci, const int* cj,
                                const double* xix, const double* xiy, const double* nx,
                                const double* ny, const double* c, const int* slot) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc) return;
    const int e = slot[k];
    t.pr_i[e] = ci[k];
    t.pr_j[e] = cj[k] | (kImgStale << kImgShift);
    t.ct_j[e] = cj[k];
    t.xix[e] = xix[k]; t.xiy[e] = xiy[k]; t.nx[e] = nx[k]; t.ny[e] = ny[k]; t.c[e] = c[k];
    t.fnx[e] = 0; t.fny[e] = 0; t.fsx[e] = 0; t.fsy[e] = 0; t.t1[e] = 0; t.t2[e] = 0;
    t.bx[e] = 0; t.by[e] = 0; t.px[e] = 0; t.py[e] = 0;
}

// ---------------------------------------------------------------------------
// host side

#define HIPCHK(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            return fail(E, XDEM_E_HIP, std::string(#call) + ": " +            \
                                           hipGetErrorString(e_));            \
        }                                                                     \
    } while (0)

}  // namespace

struct xdem_engine {
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<xdem_config> cfg;
    std::vector<Pack> packs_h;
    Pack* packs_d = nullptr;
    int n = 0;
    int npacks = 0;
    std::vector<int> pack_start;
    std::vector<int8_t> kind_h;
    std::vector<int> up_h;
    // device memory
    std::vector<void*> allocs;
    Parts P{};
    Bins B{};
    Table T[2]{};
    int cur = 0;                  // index of the table holding the last step
    uint32_t* err_d = nullptr;
    Counters* cnt_d = nullptr;
    double* stress_d = nullptr;
    long long steps = 0;
    bool have_particles = false;
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    double timed_ms = 0.0;
    long long timed_launches = 0;
    int pair_grid = 0;
};

namespace {

int fail(xdem_engine* E, int code, const std::string& msg) {
    if (E) E->err = msg;
    return code;
}

template <class T>
int dalloc(xdem_engine* E, T** p, size_t count) {
    void* q = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess)
        return fail(E, XDEM_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    E->allocs.push_back(q);
    *p = (T*)q;
    return XDEM_OK;
}

int map_error_bits(xdem_engine* E, uint32_t bits) {
    if (!bits) return XDEM_OK;
    std::string m = "step failed:";
    int code = XDEM_E_CAPACITY;
    if (bits & ERR_UNSUPPORTED) { m += " unlike particle shapes in a candidate pair (reference exit(-1), src/EllipticParticle.cpp:30);"; code = XDEM_E_UNSUPPORTED_PAIR; }
    if (bits & ERR_ZERO_NORMAL) { m += " zero-length contact normal (reference exit(-1), src/MathVector.h:187);"; code = XDEM_E_ZERO_NORMAL; }
    if (bits & ERR_NEWTON_CAP) { m += " quartic Newton reached its cap without a bracketed rescue;"; code = XDEM_E_NEWTON_CAP; }
    if (bits & ERR_CAND_OVERFLOW) { m += " a particle has >= 200 candidates (reference exit(-2));"; code = XDEM_E_CAPACITY; }
    if (bits & ERR_BIN_OVERFLOW) { m += " cell bin capacity exceeded;"; code = XDEM_E_CAPACITY; }
    if (bits & ERR_PAIR_OVERFLOW) { m += " contact table capacity exceeded;"; code = XDEM_E_CAPACITY; }
    if (bits & ERR_INCOMING_OVERFLOW) { m += " more than 32 contacts held by partners of one particle;"; code = XDEM_E_CAPACITY; }
    if (bits & ERR_ROW_OVERFLOW) { m += " more than 32 bounding-circle pairs for one particle;"; code = XDEM_E_CAPACITY; }
    return fail(E, code, m);
}

int check_error(xdem_engine* E) {
    uint32_t bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, E->err_d, sizeof(bits), hipMemcpyDeviceToHost, E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
    return map_error_bits(E, bits);
}

int alloc_table(xdem_engine* E, Table& t, int cap, int n) {
    int rc;
    t.cap = cap;
    if ((rc = dalloc(E, &t.pr_i, cap)) || (rc = dalloc(E, &t.pr_j, cap)) ||
        (rc = dalloc(E, &t.ct_j, cap)))
        return rc;
    double** d[] = {&t.xix, &t.xiy, &t.nx, &t.ny, &t.c, &t.fnx, &t.fny, &t.fsx,
                    &t.fsy, &t.t1, &t.t2, &t.bx, &t.by, &t.px, &t.py};
    for (double** q : d)
        if ((rc = dalloc(E, q, cap))) return rc;
    if ((rc = dalloc(E, &t.row_start, n)) || (rc = dalloc(E, &t.row_cnt, n)) ||
        (rc = dalloc(E, &t.counter, 1)))
        return rc;
    HIPCHK(hipMemset(t.row_start, 0, sizeof(int) * (size_t)n));
    HIPCHK(hipMemset(t.row_cnt, 0, sizeof(int) * (size_t)n));
    HIPCHK(hipMemset(t.counter, 0, sizeof(int)));
    return XDEM_OK;
}

int grid_of(int n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

// ===========================================================================
extern "C" {

int xdem_create(xdem_engine** out, int num_packings, const xdem_config* cfgs) {
    if (!out || num_packings <= 0 || !cfgs) return XDEM_E_INVALID;
    *out = nullptr;
    xdem_engine* E = new xdem_engine();
    E->npacks = num_packings;
    E->cfg.assign(cfgs, cfgs + num_packings);
    for (int k = 0; k < num_packings; ++k) {
        const xdem_config& c = cfgs[k];
        if (c.precision != cfgs[0].precision || c.precision < 0 || c.precision > 2 ||
            !(c.xmax > c.xmin) || !(c.ymax > c.ymin) || !(c.dt > 0) ||
            (c.world_kind != XDEM_WORLD_PERIODIC && c.world_kind != XDEM_WORLD_LEES_EDWARDS)) {
            delete E;
            return XDEM_E_INVALID;
        }
    }
    hipError_t he = hipGetDevice(&E->device);
    if (he == hipSuccess) he = hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        delete E;
        return XDEM_E_HIP;
    }
    int cus = 256;
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, E->device);
    E->pair_grid = cus * 8;
    *out = E;
    return XDEM_OK;
}

void xdem_destroy(xdem_engine* E) {
    if (!E) return;
    if (E->stream) hipStreamSynchronize(E->stream);
    for (void* p : E->allocs) hipFree(p);
    for (hipEvent_t ev : E->ev_pool) hipEventDestroy(ev);
    if (E->stream) hipStreamDestroy(E->stream);
    delete E;
}

const char* xdem_last_error(const xdem_engine* E) { return E ? E->err.c_str() : "null engine"; }

int xdem_set_particles(xdem_engine* E, int n, const int* packing_start, const int* kind,
                       const int* up_level, const int* flags, const double* outer_r,
                       const double* aspect, const double* mass, const double* inertia,
                       const double* x, const double* y, const double* theta,
                       const double* vx, const double* vy, const double* omega,
                       const double* clump_bx, const double* clump_by) {
    if (!E) return XDEM_E_INVALID;
    if (E->have_particles) return fail(E, XDEM_E_INVALID, "particles already set");
    if (n <= 0 || n >= (1 << kImgShift) || !kind || !up_level || !flags || !outer_r ||
        !aspect || !mass || !inertia || !x || !y || !theta || !vx || !vy || !omega)
        return fail(E, XDEM_E_INVALID, "bad particle arrays");
    std::vector<int> ps(E->npacks + 1);
    if (packing_start) ps.assign(packing_start, packing_start + E->npacks + 1);
    else if (E->npacks == 1) { ps[0] = 0; ps[1] = n; }
    else return fail(E, XDEM_E_INVALID, "packing_start required for several packings");
    if (ps[0] != 0 || ps[E->npacks] != n) return fail(E, XDEM_E_INVALID, "bad packing_start");
    for (int k = 0; k < E->npacks; ++k)
        if (ps[k + 1] <= ps[k]) return fail(E, XDEM_E_INVALID, "empty packing");
    E->n = n;
    E->pack_start = ps;
    // per-packing grid (src/PeriodicBoundaryWorld.cpp:15-22) and parameters
    std::vector<int> pack_of(n);
    int bins_total = 0;
    double max_ratio = 0.0;
    E->packs_h.resize(E->npacks);
    for (int k = 0; k < E->npacks; ++k) {
        const xdem_config& c = E->cfg[k];
        Pack& K = E->packs_h[k];
        memset(&K, 0, sizeof K);
        K.xmin = c.xmin; K.xmax = c.xmax; K.ymin = c.ymin; K.ymax = c.ymax;
        K.xlen = K.xmax - K.xmin; K.ylen = K.ymax - K.ymin;
        K.dt = c.dt; K.kn = c.kn; K.kt = c.kt; K.mu = c.mu; K.gdamp = c.global_damping;
        const double le = std::log(c.restitution);
        K.factor = 2.0 * std::sqrt(c.kn) * std::fabs(le) / std::sqrt(M_PI * M_PI + le * le);
        K.world = c.world_kind;
        K.rate = c.world_kind == XDEM_WORLD_LEES_EDWARDS ? c.shear_rate : 0.0;
        K.shift = c.total_shift;
        K.shift_vel = K.rate * K.ylen;
        K.strain = c.total_strain;
        K.newton_cap = c.newton_cap > 0 ? c.newton_cap : 200;
        K.strict = c.strict_newton;
        K.p_begin = ps[k]; K.p_end = ps[k + 1];
        double rmax = 0.0, amin = 1e300;
        for (int p = ps[k]; p < ps[k + 1]; ++p) {
            pack_of[p] = k;
            if (outer_r[p] > rmax) rmax = outer_r[p];
            if (kind[p] != XDEM_KIND_CLUMP) {
                const double area = M_PI * outer_r[p] * outer_r[p] / aspect[p];
                if (area < amin) amin = area;
            }
        }
        const double cell = c.cell_size > 0 ? c.cell_size : rmax * 2.01;
        K.nx = (int)std::floor(K.xlen / cell);
        K.ny = (int)std::floor(K.ylen / cell);
        if (K.nx < 3 || K.ny < 3)
            return fail(E, XDEM_E_GRID, "packing " + std::to_string(k) + ": cell grid smaller than 3x3");
        K.lenx = K.xlen / K.nx;
        K.leny = K.ylen / K.ny;
        K.hx = K.xlen / 2.0; K.hy = K.ylen / 2.0;
        K.bin_base = bins_total;
        bins_total += (K.ny + 2) * K.nx;
        if (amin < 1e300) max_ratio = std::max(max_ratio, K.lenx * K.leny / amin);
    }
    int bin_cap = 32;
    while (bin_cap < 2.5 * max_ratio + 8 && bin_cap < 1024) bin_cap *= 2;
    int rc;
    Parts& P = E->P;
    P.n = n;
    double** st[] = {&P.x, &P.y, &P.th, &P.vx, &P.vy, &P.w, &P.cbx, &P.cby, &P.fx, &P.fy, &P.tq};
    for (double** q : st)
        if ((rc = dalloc(E, q, n))) return rc;
    double *r_d, *ar_d, *m_d, *im_d, *iin_d;
    int *up_d, *pack_d;
    int8_t *kind_d, *flags_d;
    if ((rc = dalloc(E, &r_d, n)) || (rc = dalloc(E, &ar_d, n)) || (rc = dalloc(E, &m_d, n)) ||
        (rc = dalloc(E, &im_d, n)) || (rc = dalloc(E, &iin_d, n)) || (rc = dalloc(E, &up_d, n)) ||
        (rc = dalloc(E, &pack_d, n)) || (rc = dalloc(E, &kind_d, n)) ||
        (rc = dalloc(E, &flags_d, n)) || (rc = dalloc(E, &P.ncont, n)) ||
        (rc = dalloc(E, &P.bin_of, n)) || (rc = dalloc(E, &P.gbin_of, n)) ||
        (rc = dalloc(E, &P.in_cnt, n)) || (rc = dalloc(E, &P.in_ref, (size_t)n * kInCap)))
        return rc;
    P.r = r_d; P.ar = ar_d; P.m = m_d; P.im = im_d; P.iin = iin_d;
    P.up = up_d; P.pack = pack_d; P.kind = kind_d; P.flags = flags_d;
    std::vector<double> im(n), iin(n), cb0(n, 0.0);
    std::vector<int8_t> k8(n), f8(n);
    for (int p = 0; p < n; ++p) {
        im[p] = 1.0 / mass[p];
        iin[p] = 1.0 / inertia[p];
        k8[p] = (int8_t)kind[p];
        f8[p] = (int8_t)flags[p];
        if (kind[p] < 0 || kind[p] > 2) return fail(E, XDEM_E_INVALID, "bad particle kind");
    }
    E->kind_h = k8;
    E->up_h.assign(up_level, up_level + n);
    const size_t db = sizeof(double) * (size_t)n, ib = sizeof(int) * (size_t)n;
    HIPCHK(hipMemcpy(P.x, x, db, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(P.y, y, db, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(P.th, theta, db, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(P.vx, vx, db, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(P.vy, vy, db, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(P.w, omega, db, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(P.cbx, clump_bx ? clump_bx : cb0.data(), db, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(P.cby, clump_by ? clump_by : cb0.data(), db, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(r_d, outer_r, db, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ar_d, aspect, db, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m_d, mass, db, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(im_d, im.data(), db, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(iin_d, iin.data(), db, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(up_d, up_level, ib, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(pack_d, pack_of.data(), ib, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(kind_d, k8.data(), n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(flags_d, f8.data(), n, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(P.fx, 0, db));
    HIPCHK(hipMemset(P.fy, 0, db));
    HIPCHK(hipMemset(P.tq, 0, db));
    HIPCHK(hipMemset(P.ncont, 0, ib));
    HIPCHK(hipMemset(P.in_cnt, 0, ib));
    // bins
    E->B.cap = bin_cap;
    E->B.nbins = bins_total;
    if ((rc = dalloc(E, &E->B.count, bins_total)) ||
        (rc = dalloc(E, &E->B.list, (size_t)bins_total * bin_cap)))
        return rc;
    HIPCHK(hipMemset(E->B.count, 0, sizeof(int) * (size_t)bins_total));
    // contact tables
    const long long tcap = std::min<long long>(12LL * n + 65536, (1LL << kImgShift) - 1);
    if ((rc = alloc_table(E, E->T[0], (int)tcap, n)) || (rc = alloc_table(E, E->T[1], (int)tcap, n)))
        return rc;
    if ((rc = dalloc(E, &E->packs_d, E->npacks)) || (rc = dalloc(E, &E->err_d, 1)) ||
        (rc = dalloc(E, &E->cnt_d, 1)) || (rc = dalloc(E, &E->stress_d, (size_t)4 * n)))
        return rc;
    HIPCHK(hipMemcpy(E->packs_d, E->packs_h.data(), sizeof(Pack) * E->npacks, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(E->err_d, 0, sizeof(uint32_t)));
    HIPCHK(hipMemset(E->cnt_d, 0, sizeof(Counters)));
    E->cur = 0;
    E->have_particles = true;
    return XDEM_OK;
}

int xdem_set_contacts(xdem_engine* E, int nc, const int* ci, const int* cj,
                      const double* xi_x, const double* xi_y, const double* n_x,
                      const double* n_y, const double* damping) {
    if (!E || !E->have_particles) return fail(E, XDEM_E_INVALID, "set particles first");
    if (nc < 0 || (nc > 0 && (!ci || !cj || !xi_x || !xi_y || !n_x || !n_y || !damping)))
        return fail(E, XDEM_E_INVALID, "bad contact arrays");
    // rows of the table that the next step reads as "previous"
    std::vector<int> order(nc), cnt(E->n, 0), start(E->n, 0), slot(nc);
    for (int k = 0; k < nc; ++k) {
        if (ci[k] < 0 || ci[k] >= E->n || cj[k] <= ci[k] || cj[k] >= E->n)
            return fail(E, XDEM_E_INVALID, "contact indices must satisfy 0 <= i < j < n");
        cnt[ci[k]]++;
    }
    int acc = 0;
    for (int p = 0; p < E->n; ++p) { start[p] = acc; acc += cnt[p]; }
    if (acc > E->T[0].cap) return fail(E, XDEM_E_CAPACITY, "too many contacts");
    std::vector<int> fill(E->n, 0);
    for (int k = 0; k < nc; ++k) slot[k] = start[ci[k]] + fill[ci[k]]++;
    Table& t = E->T[E->cur];
    int rc;
    int *ci_d, *cj_d, *slot_d;
    double *a_d, *b_d, *c_d, *d_d, *e_d;
    const size_t base = E->allocs.size();
    if ((rc = dalloc(E, &ci_d, nc)) || (rc = dalloc(E, &cj_d, nc)) || (rc = dalloc(E, &slot_d, nc)) ||
        (rc = dalloc(E, &a_d, nc)) || (rc = dalloc(E, &b_d, nc)) || (rc = dalloc(E, &c_d, nc)) ||
        (rc = dalloc(E, &d_d, nc)) || (rc = dalloc(E, &e_d, nc)))
        return rc;
    if (nc > 0) {
        HIPCHK(hipMemcpy(ci_d, ci, sizeof(int) * nc, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(cj_d, cj, sizeof(int) * nc, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(slot_d, slot.data(), sizeof(int) * nc, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(a_d, xi_x, sizeof(double) * nc, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b_d, xi_y, sizeof(double) * nc, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c_d, n_x, sizeof(double) * nc, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_d, n_y, sizeof(double) * nc, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(e_d, damping, sizeof(double) * nc, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_load_contacts, dim3((nc + 255) / 256), dim3(256), 0, E->stream, t,
                           nc, ci_d, cj_d, a_d, b_d, c_d, d_d, e_d, slot_d);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipMemcpyAsync(t.row_start, start.data(), sizeof(int) * E->n, hipMemcpyHostToDevice, E->stream));
    HIPCHK(hipMemcpyAsync(t.row_cnt, cnt.data(), sizeof(int) * E->n, hipMemcpyHostToDevice, E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
    for (size_t k = base; k < E->allocs.size(); ++k) hipFree(E->allocs[k]);
    E->allocs.resize(base);
    return XDEM_OK;
}

int xdem_step(xdem_engine* E, int nsteps) {
    if (!E || !E->have_particles) return fail(E, XDEM_E_INVALID, "set particles first");
    if (nsteps < 0) return fail(E, XDEM_E_INVALID, "nsteps < 0");
    const int gn = grid_of(E->n);
    for (int s = 0; s < nsteps; ++s) {
        Table& old_t = E->T[E->cur];
        Table& new_t = E->T[E->cur ^ 1];
        hipLaunchKernelGGL(k_bin, dim3(gn), dim3(kBlock), 0, E->stream, E->P, E->packs_d, E->B, E->err_d);
        hipLaunchKernelGGL(k_candidates, dim3(gn), dim3(kBlock), 0, E->stream, E->P, E->packs_d,
                           E->B, old_t, new_t, E->err_d, E->cnt_d);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (E->timing) {
            while (E->ev_pool.size() < E->ev_used + 2) {
                hipEvent_t ev;
                HIPCHK(hipEventCreate(&ev));
                E->ev_pool.push_back(ev);
            }
            e0 = E->ev_pool[E->ev_used++];
            e1 = E->ev_pool[E->ev_used++];
            HIPCHK(hipEventRecord(e0, E->stream));
        }
        hipLaunchKernelGGL(k_pairs, dim3(E->pair_grid), dim3(kBlock), 0, E->stream, E->P,
                           E->packs_d, old_t, new_t, E->err_d, E->cnt_d);
        if (E->timing) HIPCHK(hipEventRecord(e1, E->stream));
        hipLaunchKernelGGL(k_gather, dim3(gn), dim3(kBlock), 0, E->stream, E->P, new_t, E->err_d);
        const int gi = std::max(gn, grid_of(E->B.nbins) / 4 + 1);
        hipLaunchKernelGGL(k_integrate, dim3(gi), dim3(kBlock), 0, E->stream, E->P, E->packs_d,
                           E->B, old_t.counter, E->err_d);
        HIPCHK(hipGetLastError());
        E->cur ^= 1;
        E->steps++;
    }
    return XDEM_OK;
}

int xdem_sync(xdem_engine* E) {
    if (!E) return XDEM_E_INVALID;
    HIPCHK(hipStreamSynchronize(E->stream));
    if (E->timing) {
        for (size_t k = 0; k + 1 < E->ev_used; k += 2) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, E->ev_pool[k], E->ev_pool[k + 1]));
            E->timed_ms += ms;
            E->timed_launches++;
        }
        E->ev_used = 0;
    }
    return check_error(E);
}

int xdem_get_particles(xdem_engine* E, double* x, double* y, double* theta, double* vx,
                       double* vy, double* omega) {
    int rc = xdem_sync(E);
    if (rc) return rc;
    const size_t db = sizeof(double) * (size_t)E->n;
    if (x) HIPCHK(hipMemcpy(x, E->P.x, db, hipMemcpyDeviceToHost));
    if (y) HIPCHK(hipMemcpy(y, E->P.y, db, hipMemcpyDeviceToHost));
    if (theta) HIPCHK(hipMemcpy(theta, E->P.th, db, hipMemcpyDeviceToHost));
    if (vx) HIPCHK(hipMemcpy(vx, E->P.vx, db, hipMemcpyDeviceToHost));
    if (vy) HIPCHK(hipMemcpy(vy, E->P.vy, db, hipMemcpyDeviceToHost));
    if (omega) HIPCHK(hipMemcpy(omega, E->P.w, db, hipMemcpyDeviceToHost));
    return XDEM_OK;
}

int xdem_get_forces(xdem_engine* E, double* fx, double* fy, double* torque, int* num_contacts) {
    int rc = xdem_sync(E);
    if (rc) return rc;
    const size_t db = sizeof(double) * (size_t)E->n;
    if (fx) HIPCHK(hipMemcpy(fx, E->P.fx, db, hipMemcpyDeviceToHost));
    if (fy) HIPCHK(hipMemcpy(fy, E->P.fy, db, hipMemcpyDeviceToHost));
    if (torque) HIPCHK(hipMemcpy(torque, E->P.tq, db, hipMemcpyDeviceToHost));
    if (num_contacts) HIPCHK(hipMemcpy(num_contacts, E->P.ncont, sizeof(int) * (size_t)E->n, hipMemcpyDeviceToHost));
    return XDEM_OK;
}

int xdem_get_contacts(xdem_engine* E, int cap, int* nc, int* ci, int* cj, double* fn_x,
                      double* fn_y, double* fs_x, double* fs_y, double* point_x,
                      double* point_y, double* xi_x, double* xi_y, double* n_x,
                      double* n_y, double* damping, double* torque_i, double* torque_j) {
    int rc = xdem_sync(E);
    if (rc) return rc;
    if (!nc) return fail(E, XDEM_E_INVALID, "nc is NULL");
    const Table& t = E->T[E->cur];
    int used = 0;
    HIPCHK(hipMemcpy(&used, t.counter, sizeof(int), hipMemcpyDeviceToHost));
    if (E->steps == 0) {          // restart table: rows only
        std::vector<int> cnt(E->n);
        HIPCHK(hipMemcpy(cnt.data(), t.row_cnt, sizeof(int) * E->n, hipMemcpyDeviceToHost));
        used = 0;
        for (int v : cnt) used += v;
    }
    used = std::min(used, t.cap);
    std::vector<int> pi(used), pj(used);
    HIPCHK(hipMemcpy(pi.data(), t.pr_i, sizeof(int) * used, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(pj.data(), t.ct_j, sizeof(int) * used, hipMemcpyDeviceToHost));
    std::vector<int> idx;
    idx.reserve(used);
    for (int e = 0; e < used; ++e)
        if (pj[e] >= 0) idx.push_back(e);
    std::sort(idx.begin(), idx.end(), [&](int a, int b) {
        return pi[a] != pi[b] ? pi[a] < pi[b] : pj[a] < pj[b];
    });
    *nc = (int)idx.size();
    const int m = std::min<int>(cap, (int)idx.size());
    std::vector<double> buf(used);
    auto fetch = [&](double* dst, const double* src) -> int {
        if (!dst) return XDEM_OK;
        HIPCHK(hipMemcpy(buf.data(), src, sizeof(double) * used, hipMemcpyDeviceToHost));
        for (int k = 0; k < m; ++k) dst[k] = buf[idx[k]];
        return XDEM_OK;
    };
    for (int k = 0; k < m; ++k) {
        if (ci) ci[k] = pi[idx[k]];
        if (cj) cj[k] = pj[idx[k]];
    }
    if ((rc = fetch(fn_x, t.fnx)) || (rc = fetch(fn_y, t.fny)) || (rc = fetch(fs_x, t.fsx)) ||
        (rc = fetch(fs_y, t.fsy)) || (rc = fetch(point_x, t.px)) || (rc = fetch(point_y, t.py)) ||
        (rc = fetch(xi_x, t.xix)) || (rc = fetch(xi_y, t.xiy)) || (rc = fetch(n_x, t.nx)) ||
        (rc = fetch(n_y, t.ny)) || (rc = fetch(damping, t.c)) || (rc = fetch(torque_i, t.t1)) ||
        (rc = fetch(torque_j, t.t2)))
        return rc;
    return XDEM_OK;
}

int xdem_get_stress(xdem_engine* E, double* out) {
    int rc = xdem_sync(E);
    if (rc) return rc;
    if (!out) return fail(E, XDEM_E_INVALID, "out is NULL");
    hipLaunchKernelGGL(k_stress, dim3(grid_of(E->n)), dim3(kBlock), 0, E->stream, E->P,
                       E->T[E->cur], E->stress_d);
    HIPCHK(hipGetLastError());
    std::vector<double> s((size_t)4 * E->n);
    HIPCHK(hipMemcpyAsync(s.data(), E->stress_d, sizeof(double) * s.size(), hipMemcpyDeviceToHost, E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
    for (int k = 0; k < E->npacks; ++k) {
        double a = 0, b = 0, c = 0, d = 0;
        for (int p = E->pack_start[k]; p < E->pack_start[k + 1]; ++p) {
            a += s[4 * (size_t)p]; b += s[4 * (size_t)p + 1];
            c += s[4 * (size_t)p + 2]; d += s[4 * (size_t)p + 3];
        }
        const Pack& K = E->packs_h[k];
        const double area = K.xlen * K.ylen;
        out[4 * k] = a / area; out[4 * k + 1] = b / area;
        out[4 * k + 2] = c / area; out[4 * k + 3] = d / area;
    }
    return XDEM_OK;
}

int xdem_get_boundary(xdem_engine* E, double* out) {
    int rc = xdem_sync(E);
    if (rc) return rc;
    std::vector<Pack> h(E->npacks);
    HIPCHK(hipMemcpy(h.data(), E->packs_d, sizeof(Pack) * E->npacks, hipMemcpyDeviceToHost));
    for (int k = 0; k < E->npacks; ++k) {
        out[3 * k] = h[k].shift; out[3 * k + 1] = h[k].shift_vel; out[3 * k + 2] = h[k].strain;
    }
    return XDEM_OK;
}

int xdem_get_counters(xdem_engine* E, long long* newton_iters, long long* rescues,
                      long long* fallbacks, long long* candidates, long long* coarse_pairs,
                      long long* steps) {
    int rc = xdem_sync(E);
    if (rc) return rc;
    Counters c;
    HIPCHK(hipMemcpy(&c, E->cnt_d, sizeof c, hipMemcpyDeviceToHost));
    if (newton_iters) *newton_iters = (long long)c.newton;
    if (rescues) *rescues = (long long)c.rescues;
    if (fallbacks) *fallbacks = (long long)c.fallbacks;
    if (candidates) *candidates = (long long)c.cand;
    if (coarse_pairs) *coarse_pairs = (long long)c.coarse;
    if (steps) *steps = E->steps;
    return XDEM_OK;
}

int xdem_set_timing(xdem_engine* E, int enable) {
    if (!E) return XDEM_E_INVALID;
    int rc = xdem_sync(E);
    if (rc) return rc;
    E->timing = enable != 0;
    E->timed_ms = 0.0;
    E->timed_launches = 0;
    return XDEM_OK;
}

int xdem_get_timing(xdem_engine* E, double* contact_ms, long long* launches) {
    int rc = xdem_sync(E);
    if (rc) return rc;
    if (contact_ms) *contact_ms = E->timed_ms;
    if (launches) *launches = E->timed_launches;
    return XDEM_OK;
}

int xdem_get_traffic(xdem_engine* E, double* bytes_per_step, long long* contacts) {
    int rc = xdem_sync(E);
    if (rc) return rc;
    const Table& t = E->T[E->cur];
    int used = 0;
    HIPCHK(hipMemcpy(&used, t.counter, sizeof(int), hipMemcpyDeviceToHost));
    used = std::min(used, t.cap);
    std::vector<int> pj(used);
    HIPCHK(hipMemcpy(pj.data(), t.ct_j, sizeof(int) * used, hipMemcpyDeviceToHost));
    long long live = 0;
    for (int v : pj) live += v >= 0;
    // state 96 B r/w + static 44 B per particle; 44 B history r/w per contact
    if (bytes_per_step) *bytes_per_step = 140.0 * E->n + 88.0 * (double)live;
    if (contacts) *contacts = live;
    return XDEM_OK;
}

int xdem_set_mlp(xdem_engine* E, const float*, const float*, const float*, const float*,
                 const float*, const float*) {
    return fail(E, XDEM_E_INVALID, "learned root guess not available in this build");
}

int xdem_log_roots(xdem_engine* E, int, int* count, float*, double*, int*) {
    if (count) *count = 0;
    return fail(E, XDEM_E_INVALID, "root logging not available in this build");
}

}  // extern "C"
