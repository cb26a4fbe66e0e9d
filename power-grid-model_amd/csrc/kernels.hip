// kernels.hip -- gfx950 kernels of the batched Newton-Raphson power flow.
//
// One wave = 64 scenarios of one group running the same task (a connected piece of the elimination
// tree). A lane walks the task's pivots in increasing order (forward sweep) or decreasing order
// (backward sweep); every per-scenario access is a coalesced 64-lane load/store because all
// per-scenario arrays are scenario-interleaved. The pattern, the admittances and the schedule are
// the same for every lane and are read through the scalar cache.
//
// Math restated from the reference (paths relative to power_grid_model_c/power_grid_model/include/
// power_grid_model/math_solver/):
//   linear initial guess     newton_raphson_pf_solver.hpp:255-303, 706-762, 856-870
//   Jacobian and mismatch    newton_raphson_pf_solver.hpp:462-547, 764-852
//   full-pivot block LU      sparse_lu_solver.hpp:85-165 (first maximum of |a|^2 in column-major
//                            order, SparseMatrixError when |u_ii| < eps * max pivot or not normal)
//   sparse LU + solves       sparse_lu_solver.hpp:346-495, 769-827 (right-looking there; here the
//                            same updates are pulled by the pivot that owns each entry, in the same
//                            increasing elimination order)
//   voltage update           newton_raphson_pf_solver.hpp:325-349
//   results                  y_bus.hpp:482-546, common_solver_functions.hpp:80-160, 383-446
#include "kernels.hpp"
#include "records.hpp"

#include <cfloat>

namespace pgm_amd {

// kernels keep external linkage so profilers and the code-object loader can resolve their symbols
namespace dev {

constexpr int kWave = 64;

// Group descriptors, work lists and structure arrays are written by the host before the launches
// that read them. Reading them through the constant address space lets the compiler use scalar loads
// where the index is the same on every lane, and a pointer field read from a descriptor this way is a
// global-memory pointer, so the per-lane data it addresses is reached with global (not flat) accesses.
// (The host pass only parses kernel bodies; it sees plain loads.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class T> using const_ptr = T const __attribute__((address_space(4)))*;
template <class T, class I> __device__ __forceinline__ T ld(T const* p, I i) { return ((const_ptr<T>)p)[i]; }
// byte arrays: the aligned 32-bit word holding the byte (scalar loads read whole dwords; every device
// array starts on a 256-byte boundary and is padded to a multiple of 256 bytes, so the word is inside it)
template <class I> __device__ __forceinline__ int8_t ld(int8_t const* p, I i) {
    uintptr_t const a = reinterpret_cast<uintptr_t>(p) + uintptr_t(i);
    uint32_t const w = ((const_ptr<uint32_t>)(a & ~uintptr_t(3)))[0];
    return int8_t(w >> ((a & 3) * 8));
}
#else
template <class T, class I> __device__ T ld(T const* p, I i) { return p[i]; }
#endif

// group of a flat thread index: offsets[g] <= idx < offsets[g + 1]
__device__ __forceinline__ int find_group(long long const* __restrict__ offsets, int n_groups, long long idx) {
    int lo = 0, hi = n_groups;
    while (hi - lo > 1) {
        int const mid = (lo + hi) >> 1;
        if (ld(offsets, mid) <= idx)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

struct cplx {
    double re, im;
};
__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cplx cconj(cplx a) { return {a.re, -a.im}; }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cplx cdiv(cplx a, cplx b) {
    double const d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
__device__ __forceinline__ double cabs_(cplx a) { return sqrt(a.re * a.re + a.im * a.im); }

template <class FT> __device__ __forceinline__ bool is_normal_(FT x) {
    FT const ax = x < FT(0) ? -x : x;
    FT const mn = sizeof(FT) == 8 ? FT(DBL_MIN) : FT(FLT_MIN);
    FT const mx = sizeof(FT) == 8 ? FT(DBL_MAX) : FT(FLT_MAX);
    return ax >= mn && ax <= mx;  // false for 0, subnormal, inf, nan
}

// ---------------- dense NB x NB blocks, column-major a[c * NB + r] ----------------
template <int NB, class T> __device__ __forceinline__ void swap_rows(T (&a)[NB * NB], int t, int r_sel) {
#pragma unroll
    for (int r = 0; r < NB; ++r) {
        if (r <= t) continue;
        bool const sw = r_sel == r;
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            T const x = a[c * NB + t], y = a[c * NB + r];
            a[c * NB + t] = sw ? y : x;
            a[c * NB + r] = sw ? x : y;
        }
    }
}
template <int NB, class T> __device__ __forceinline__ void swap_cols(T (&a)[NB * NB], int t, int c_sel) {
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        if (c <= t) continue;
        bool const sw = c_sel == c;
#pragma unroll
        for (int r = 0; r < NB; ++r) {
            T const x = a[t * NB + r], y = a[c * NB + r];
            a[t * NB + r] = sw ? y : x;
            a[c * NB + r] = sw ? x : y;
        }
    }
}
__device__ __forceinline__ int perm_row(uint32_t code, int t) { return int((code >> (3 * t)) & 7u); }
__device__ __forceinline__ int perm_col(uint32_t code, int t) { return int((code >> (15 + 3 * t)) & 7u); }

// full-pivot LU in place: P a Q = L U; returns false when the reference would raise SparseMatrixError
// perturb_thr > 0 (perturb_pivot_if_needed, sparse_lu_solver.hpp:39-48 and :120-124): a chosen pivot whose
// magnitude is below the threshold takes the threshold's magnitude with its own sign (+ for a zero pivot) and
// `perturbed` is set; the pivot test then checks only for non-normal values, as the reference's does once it
// has perturbed (:157-164)
// skip_relative: an earlier pivot of the same system was perturbed (the reference's flag is per solver, so from
// then on only non-normal pivots fail)
template <int NB, class FT>
__device__ __forceinline__ bool factor_block(FT (&a)[NB * NB], uint32_t& code, FT perturb_thr, bool& perturbed,
                                             bool skip_relative = false) {
    FT max_pivot = FT(0);
    bool zero = false;
    bool pert = false;
    code = 0;
#pragma unroll
    for (int t = 0; t < NB; ++t) {
        FT best = FT(-1);
        int rb = t, cb = t;
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            if (c < t) continue;
#pragma unroll
            for (int r = 0; r < NB; ++r) {
                if (r < t) continue;
                FT const v = a[c * NB + r] * a[c * NB + r];
                bool const gt = v > best;
                best = gt ? v : best;
                rb = gt ? r : rb;
                cb = gt ? c : cb;
            }
        }
        if (t < NB - 1) {
            code |= uint32_t(rb) << (3 * t);
            code |= uint32_t(cb) << (15 + 3 * t);
            swap_rows<NB, FT>(a, t, rb);
            swap_cols<NB, FT>(a, t, cb);
        }
        // the chosen pivot now sits at (t, t)
        if (perturb_thr > FT(0) && best < perturb_thr * perturb_thr) {
            a[t * NB + t] = a[t * NB + t] < FT(0) ? -perturb_thr : perturb_thr;
            best = perturb_thr * perturb_thr;
            pert = true;
        }
        zero = zero || best == FT(0);
        max_pivot = sqrt(best) > max_pivot ? sqrt(best) : max_pivot;
        if (t < NB - 1) {
            FT const piv = a[t * NB + t];
#pragma unroll
            for (int r = 0; r < NB; ++r)
                if (r > t) a[t * NB + r] /= piv;
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                if (c <= t) continue;
#pragma unroll
                for (int r = 0; r < NB; ++r)
                    if (r > t) a[c * NB + r] -= a[t * NB + r] * a[c * NB + t];
            }
        }
    }
    FT const eps = sizeof(FT) == 8 ? FT(DBL_EPSILON) : FT(FLT_EPSILON);
    FT const thr = (pert || skip_relative) ? FT(0) : eps * max_pivot;
    perturbed = pert;
    bool ok = !zero;
#pragma unroll
    for (int t = 0; t < NB; ++t) {
        FT const d = a[t * NB + t];
        FT const ad = d < FT(0) ? -d : d;
        ok = ok && !(ad < thr) && is_normal_(d);
    }
    return ok;
}
template <int NB, class FT> __device__ __forceinline__ bool factor_block(FT (&a)[NB * NB], uint32_t& code) {
    bool unused = false;
    return factor_block<NB, FT>(a, code, FT(0), unused);
}
template <int NB, class T> __device__ __forceinline__ void apply_p_rows(T (&m)[NB * NB], uint32_t code) {
#pragma unroll
    for (int t = 0; t < NB - 1; ++t) swap_rows<NB, T>(m, t, perm_row(code, t));
}
template <int NB, class T> __device__ __forceinline__ void apply_q_cols(T (&m)[NB * NB], uint32_t code) {
#pragma unroll
    for (int t = 0; t < NB - 1; ++t) swap_cols<NB, T>(m, t, perm_col(code, t));
}
template <int NB> __device__ __forceinline__ void apply_p_vec(double (&v)[NB], uint32_t code) {
#pragma unroll
    for (int t = 0; t < NB - 1; ++t) {
        int const r = perm_row(code, t);
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            if (k <= t) continue;
            bool const sw = r == k;
            double const x = v[t], y = v[k];
            v[t] = sw ? y : x;
            v[k] = sw ? x : y;
        }
    }
}
template <int NB> __device__ __forceinline__ void apply_q_vec(double (&v)[NB], uint32_t code) {
#pragma unroll
    for (int t = NB - 2; t >= 0; --t) {
        int const c = perm_col(code, t);
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            if (k <= t) continue;
            bool const sw = c == k;
            double const x = v[t], y = v[k];
            v[t] = sw ? y : x;
            v[k] = sw ? x : y;
        }
    }
}
// m := L^-1 m, L unit lower in lu
template <int NB, class FT> __device__ __forceinline__ void lower_left(FT const (&lu)[NB * NB], FT (&m)[NB * NB]) {
#pragma unroll
    for (int r = 1; r < NB; ++r)
#pragma unroll
        for (int k = 0; k < r; ++k) {
            FT const l = lu[k * NB + r];
#pragma unroll
            for (int c = 0; c < NB; ++c) m[c * NB + r] -= l * m[c * NB + k];
        }
}
// m := m U^-1
template <int NB, class FT> __device__ __forceinline__ void upper_right(FT const (&lu)[NB * NB], FT (&m)[NB * NB]) {
#pragma unroll
    for (int c = 0; c < NB; ++c) {
#pragma unroll
        for (int k = 0; k < c; ++k) {
            FT const u = lu[c * NB + k];
#pragma unroll
            for (int r = 0; r < NB; ++r) m[c * NB + r] -= u * m[k * NB + r];
        }
        FT const d = lu[c * NB + c];
#pragma unroll
        for (int r = 0; r < NB; ++r) m[c * NB + r] /= d;
    }
}
template <int NB, class FT> __device__ __forceinline__ void lower_vec(FT const (&lu)[NB * NB], double (&v)[NB]) {
#pragma unroll
    for (int r = 1; r < NB; ++r)
#pragma unroll
        for (int k = 0; k < r; ++k) v[r] -= double(lu[k * NB + r]) * v[k];
}
template <int NB, class FT> __device__ __forceinline__ void upper_vec(FT const (&lu)[NB * NB], double (&v)[NB]) {
#pragma unroll
    for (int r = NB - 1; r >= 0; --r) {
#pragma unroll
        for (int k = NB - 1; k > r; --k) v[r] -= double(lu[k * NB + r]) * v[k];
        v[r] /= double(lu[r * NB + r]);
    }
}

template <int NB, class FT> struct Blocks {
    FT* base;
    size_t sp;
    int s;
    __device__ __forceinline__ FT* ptr(int e, int k) const { return base + ((size_t(e) * (NB * NB) + k) * sp + s); }
    __device__ __forceinline__ void load(int e, FT (&m)[NB * NB]) const {
#pragma unroll
        for (int k = 0; k < NB * NB; ++k) m[k] = *ptr(e, k);
    }
    __device__ __forceinline__ void store(int e, FT const (&m)[NB * NB]) const {
#pragma unroll
        for (int k = 0; k < NB * NB; ++k) *ptr(e, k) = m[k];
    }
    __device__ __forceinline__ void store_zero(int e) const {
#pragma unroll
        for (int k = 0; k < NB * NB; ++k) *ptr(e, k) = FT(0);
    }
};

// HNML block of entry (i, j): pf = (u_i outer conj u_j) .* conj(Y_ij); H = L = Im pf, N = -M = Re pf
template <int P, class T>
__device__ __forceinline__ void hnml(double const* y, cplx const (&ui)[P], cplx const (&uj)[P], T (&b)[4 * P * P],
                                     double (&prow)[P], double (&qrow)[P]) {
    constexpr int NB = 2 * P;
#pragma unroll
    for (int r = 0; r < P; ++r) {
        prow[r] = 0.0;
        qrow[r] = 0.0;
#pragma unroll
        for (int c = 0; c < P; ++c) {
            cplx const yc = {y[(r * P + c) * 2], -y[(r * P + c) * 2 + 1]};
            cplx const pf = cmul(cmul(ui[r], cconj(uj[c])), yc);
            b[c * NB + r] = T(pf.im);
            b[(c + P) * NB + r] = T(pf.re);
            b[c * NB + r + P] = T(-pf.re);
            b[(c + P) * NB + r + P] = T(pf.im);
            prow[r] += pf.re;
            qrow[r] += pf.im;
        }
    }
}
// [[G, -B], [B, G]] of an admittance tensor
template <int P, class T> __device__ __forceinline__ void linear_block(double const* y, T (&b)[4 * P * P]) {
    constexpr int NB = 2 * P;
#pragma unroll
    for (int r = 0; r < P; ++r)
#pragma unroll
        for (int c = 0; c < P; ++c) {
            double const g = y[(r * P + c) * 2], bb = y[(r * P + c) * 2 + 1];
            b[c * NB + r] = T(g);
            b[(c + P) * NB + r] = T(-bb);
            b[c * NB + r + P] = T(bb);
            b[(c + P) * NB + r + P] = T(g);
        }
}
// admittance tensor of a source: y1 (sym) or (2 y1 + y0) / 3 on, (y0 - y1) / 3 off the diagonal
template <int P> __device__ __forceinline__ void source_y(cplx y1, cplx y0, double (&y)[P * P * 2]) {
#pragma unroll
    for (int r = 0; r < P; ++r)
#pragma unroll
        for (int c = 0; c < P; ++c) {
            cplx v;
            if (P == 1)
                v = y1;
            else if (r == c)
                v = {(2.0 * y1.re + y0.re) / 3.0, (2.0 * y1.im + y0.im) / 3.0};
            else
                v = {(y0.re - y1.re) / 3.0, (y0.im - y1.im) / 3.0};
            y[(r * P + c) * 2] = v.re;
            y[(r * P + c) * 2 + 1] = v.im;
        }
}
template <int P> __device__ __forceinline__ void balanced(cplx x, cplx (&v)[P]) {
    v[0] = x;
    if (P == 3) {
        cplx const a2 = {-0.5, -0.86602540378443864676};
        cplx const a = {-0.5, 0.86602540378443864676};
        v[1 % P] = cmul(x, a2);
        v[2 % P] = cmul(x, a);
    }
}

template <int P> __device__ __forceinline__ void load_u(DevGroup const& g, int bus, int s, cplx (&u)[P]) {
    size_t const sp = g.S_pad;
#pragma unroll
    for (int ph = 0; ph < P; ++ph) {
        u[ph].re = g.u[((size_t(bus) * P + ph) * 2) * sp + s];
        u[ph].im = g.u[((size_t(bus) * P + ph) * 2 + 1) * sp + s];
    }
}

// Y values of LU entry k for lane s: shared by the group, or per lane in union groups
template <int P, bool UNI>
__device__ __forceinline__ void load_y(DevGroup const& g, int k, int s, double (&y)[P * P * 2]) {
    constexpr int Q = P * P * 2;
    if constexpr (UNI) {
        size_t const sp = g.S_pad;
#pragma unroll
        for (int q = 0; q < Q; ++q) y[q] = g.ylane[(size_t(k) * Q + q) * sp + s];
    } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) y[q] = ld(g.ydata, size_t(k) * Q + q);
    }
}

// task and lane of a thread. Plain and union groups: one task, 64 consecutive scenarios per wave.
// Packed groups (at most 32 scenarios): a wave holds 64 / S tasks of one level, S lanes each
// (item.task = first index into the level's task list, item.block = number of tasks)
template <int KIND>
__device__ __forceinline__ bool lane_work(DevGroup const& g, WorkItem const& wi, int& task, int& s) {
    if constexpr (KIND == KIND_PACKED) {
        int const t = int(threadIdx.x) / g.S;
        s = int(threadIdx.x) - t * g.S;
        if (t >= wi.block) return false;
        task = ld(g.level_task, wi.task + t);
        return true;
    } else {
        task = wi.task;
        s = wi.block * kWave + int(threadIdx.x);
        return s < g.S;
    }
}

// extra (off-pattern) entries of row p in lane s of a union group (branches closed in the lane that
// the shared structure has open): MODE_INIT keeps the admittance block, MODE_NR the Jacobian block and
// adds the entry's power terms to the row sums. The blocks enter through the low-rank correction.
template <int MODE>
__device__ __forceinline__ void extra_entries(DevGroup const& g, int p, int s, cplx const (&up)[1], double& pcal,
                                              double& qcal) {
    int const nx = ld(g.xcnt, s);
    for (int e = 0; e < nx; ++e) {
        size_t const q = size_t(s) * kExtraMax + e;
        if (ld(g.xrow, q) != p) continue;
        double const y[2] = {ld(g.xy, 2 * q), ld(g.xy, 2 * q + 1)};
        double blk[4];
        if constexpr (MODE == MODE_INIT) {
            linear_block<1, double>(y, blk);
        } else {
            cplx uj[1];
            load_u<1>(g, ld(g.xcol, q), s, uj);
            double pr[1], qr[1];
            hnml<1, double>(y, up, uj, blk, pr, qr);
            pcal += pr[0];
            qcal += qr[0];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) g.xblk[q * 4 + k] = blk[k];
    }
}

// ---------------------------------------------------------------------------------------------------
// forward sweep: assemble row / column of each pivot, pull the Schur updates, factor the diagonal
// block, finish the U row and L column, forward-substitute
// ---------------------------------------------------------------------------------------------------
// waves per SIMD the sweeps are compiled for where 80 VGPRs hold without spilling
// the bus's own terms of the Newton-Raphson system once the branch sums pcal, qcal of row p are formed: the
// mismatch (rhs = -P_cal, -Q_cal, plus the loads) and the diagonal block's power, load (ZIP) and source terms
// (newton_raphson_pf_solver.hpp:473-547, 764-852)
template <int P>
__device__ __forceinline__ void nr_bus_terms(DevGroup const& g, int p, int s, cplx const (&up)[P], double const (&pcal)[P],
                                             double const (&qcal)[P], double (&D)[4 * P * P], double (&rhs)[2 * P]) {
    constexpr int NB = 2 * P;
    constexpr int NB2 = NB * NB;
    constexpr int PP = P * P;
    size_t const sp = g.S_pad;
#pragma unroll
    for (int ph = 0; ph < P; ++ph) {
        rhs[ph] = -pcal[ph];
        rhs[ph + P] = -qcal[ph];
        D[ph * NB + ph] += -qcal[ph];
        D[(ph + P) * NB + ph] += pcal[ph];
        D[ph * NB + ph + P] += pcal[ph];
        D[(ph + P) * NB + ph + P] += qcal[ph];
    }
    // loads: constant power / impedance / current (newton_raphson_pf_solver.hpp:764-822)
    for (int l = ld(g.load_ptr, p); l < ld(g.load_ptr, p + 1); ++l) {
        int const type = ld(g.load_type, l);
#pragma unroll
        for (int ph = 0; ph < P; ++ph) {
            double const sr = g.s_inj[((size_t(l) * P + ph) * 2) * sp + s];
            double const si = g.s_inj[((size_t(l) * P + ph) * 2 + 1) * sp + s];
            double const v = g.xv[((size_t(p) * P + ph) * 2 + 1) * sp + s];
            double const f = type == 0 ? 1.0 : (type == 1 ? v * v : v);
            rhs[ph] += sr * f;
            rhs[ph + P] += si * f;
            double const df = type == 0 ? 0.0 : (type == 1 ? 2.0 * v * v : v);
            D[(ph + P) * NB + ph] += -sr * df;
            D[(ph + P) * NB + ph + P] += -si * df;
        }
    }
    // sources: two-bus equivalent with y_ref (newton_raphson_pf_solver.hpp:824-852)
    for (int q = ld(g.src_ptr, p); q < ld(g.src_ptr, p + 1); ++q) {
        cplx const y1 = {g.y_src[(size_t(q) * 4) * sp + s], g.y_src[(size_t(q) * 4 + 1) * sp + s]};
        cplx const y0 = {g.y_src[(size_t(q) * 4 + 2) * sp + s], g.y_src[(size_t(q) * 4 + 3) * sp + s]};
        cplx const ur = {g.u_src[(size_t(q) * 2) * sp + s], g.u_src[(size_t(q) * 2 + 1) * sp + s]};
        double ys[PP * 2], yn[PP * 2];
        source_y<P>(y1, y0, ys);
#pragma unroll
        for (int k = 0; k < PP * 2; ++k) yn[k] = -ys[k];
        cplx ub[P];
        balanced<P>(ur, ub);
        double bmm[NB2], bms[NB2];
        double p1[P], q1[P], p2[P], q2[P];
        hnml<P, double>(ys, up, up, bmm, p1, q1);
        hnml<P, double>(yn, up, ub, bms, p2, q2);
#pragma unroll
        for (int r = 0; r < P; ++r) {
            double const pc = p1[r] + p2[r];
            double const qc = q1[r] + q2[r];
            bmm[r * NB + r] += -qc;
            bmm[(r + P) * NB + r] += pc;
            bmm[r * NB + r + P] += pc;
            bmm[(r + P) * NB + r + P] += qc;
            rhs[r] -= pc;
            rhs[r + P] -= qc;
        }
#pragma unroll
        for (int k = 0; k < NB2; ++k) D[k] += bmm[k];
    }
}

template <int P, class FT, int KIND> constexpr int forward_waves() {
    return P == 1 && KIND == KIND_PLAIN && sizeof(FT) == 4 ? 5 : 1;
}
template <int P, int MODE, int KIND> constexpr int backward_waves() {
    return P == 1 && KIND != KIND_PACKED ? (MODE == MODE_NR ? 8 : 6) : 1;
}

template <int P, class FT, int MODE, int KIND>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(forward_waves<P, FT, KIND>()))) void forward_kernel(DevGroup const* __restrict__ groups,
                                                        WorkItem const* __restrict__ items) {
    constexpr int NB = 2 * P;
    constexpr int NB2 = NB * NB;
    constexpr int PP = P * P;
    constexpr bool UNI = KIND == KIND_UNION;
    WorkItem const wi = ld(items, blockIdx.x);
    DevGroup const g = ld(groups, wi.group);
    int task, s;
    if (!lane_work<KIND>(g, wi, task, s)) return;
    size_t const sp = g.S_pad;
    int const mm = ld(g.task_mm, task);
    if (g.done[size_t(mm) * sp + s] != (MODE == MODE_NRA ? CONVERGED : MODE == MODE_REFINE ? REFINE : ACTIVE)) return;
    Blocks<NB, FT> const lu{static_cast<FT*>(g.lu), sp, s};
    // a keep row whose parent is the next pivot of the task hands its finished U_p,parent,
    // L_parent,p and y_p to that pivot through LDS (fw_le: label of the forwarded L entry); LDS keeps
    // the hand-over off the vector registers
    constexpr bool kFwd = P == 1;
    __shared__ FT fw_blk[kFwd ? kWave : 1][kFwd ? 2 * NB2 : 1];
    __shared__ double fw_y[kFwd ? kWave : 1][kFwd ? NB : 1];
    // Newton-Raphson: the handing row's own voltage and its parent's, which the parent reads again
    __shared__ double fw_u[kFwd ? kWave : 1][4];
    int const lane = int(threadIdx.x);
    int fw_le = -1;
    bool fw_uok = false;  // fw_u holds the last row's voltage and its parent's
    // L_parent,p of a handing row is needed by nobody but the parent, which takes it from LDS
    // (NR and the linear system; union groups keep it for the low-rank correction's chains)
    constexpr bool kSkipL = kFwd && (MODE == MODE_NR || MODE == MODE_INIT) && KIND != KIND_UNION;
    // FP32 factors of a Newton-Raphson sweep in plain and packed groups may perturb small pivots
    constexpr bool kPerturb = sizeof(FT) == 4 && MODE == MODE_NR && KIND != KIND_UNION;
    __shared__ FT thr_s[kPerturb ? kWave : 1];
    int handed = -1;  // MODE_REFINE: the previous pivot when it handed its L block over (not stored)
    // the state estimation's gain systems (MODE_RAWP): a pivot of this pair was perturbed in an earlier launch
    // (read once; launches are ordered) or earlier in this task
    bool raw_pert = false;
    if constexpr (MODE == MODE_RAWP) raw_pert = g.pthr && g.pert && g.pert[size_t(mm) * sp + s] != 0;

    for (int ti = ld(g.task_ptr, task); ti < ld(g.task_ptr, task + 1); ++ti) {
        int const p = ld(g.task_pivot, ti);
        int const rb = ld(g.row_ptr, p), re = ld(g.row_end, p), dg = ld(g.diag, p);
        int const es = g.row_stride ? ld(g.row_stride, p) : 1;
        if constexpr (MODE == MODE_REFINE) {
            // forward substitution of the residual (in y) with the stored factors: y_p = L_pp^-1 P (r_p -
            // sum_k L_pk y_k). A child that handed its results to p through LDS in the Newton-Raphson sweep did
            // not store its L block: it is formed again as that sweep formed it, J_pj Q_j U_jj^-1 (a keep row
            // takes no Schur update in that block)
            double v[NB];
#pragma unroll
            for (int r = 0; r < NB; ++r) v[r] = g.y[(size_t(p) * NB + r) * sp + s];
            for (int k = rb; k < dg; k += es) {
                int const j = ld(g.col, k);
                FT L[NB2];
                if (kFwd && KIND != KIND_UNION && j == handed) {
                    if (ld(g.fill, k)) {
#pragma unroll
                        for (int q = 0; q < NB2; ++q) L[q] = FT(0);
                    } else {
                        cplx up[P], uj[P];
                        load_u<P>(g, p, s, up);
                        load_u<P>(g, j, s, uj);
                        double yk[PP * 2], pr[P], qr[P];
                        load_y<P, false>(g, k, s, yk);
                        hnml<P, FT>(yk, up, uj, L, pr, qr);
                        FT Dj[NB2];
                        lu.load(ld(g.diag, j), Dj);
                        apply_q_cols<NB, FT>(L, g.perm[size_t(j) * sp + s]);
                        upper_right<NB, FT>(Dj, L);
                    }
                } else {
                    lu.load(k, L);
                }
                double yj[NB];
#pragma unroll
                for (int r = 0; r < NB; ++r) yj[r] = g.y[(size_t(j) * NB + r) * sp + s];
#pragma unroll
                for (int r = 0; r < NB; ++r) {
                    double acc = 0.0;
#pragma unroll
                    for (int c = 0; c < NB; ++c) acc += double(L[c * NB + r]) * yj[c];
                    v[r] -= acc;
                }
            }
            FT Dst[NB2];
            lu.load(dg, Dst);
            apply_p_vec<NB>(v, g.perm[size_t(p) * sp + s]);
            lower_vec<NB, FT>(Dst, v);
#pragma unroll
            for (int r = 0; r < NB; ++r) g.y[(size_t(p) * NB + r) * sp + s] = v[r];
            // the hand-over condition of the Newton-Raphson sweep (keep row whose parent is the next pivot)
            bool const keep = P == 1 && ld(g.keep, p) != 0;
            handed = keep && ti + 1 < ld(g.task_ptr, task + 1) && ld(g.task_pivot, ti + 1) == ld(g.col, dg + es)
                         ? p
                         : -1;
            continue;
        }
        // a bus without a source in this lane (union groups): identity row, zero row and column
        bool const dead_p = UNI && g.dead[size_t(p) * sp + s] != 0;
        // rows with one U entry and no off-diagonal Schur target (every bus of a radial feeder) keep
        // their raw U entry and its transpose in registers until the row is finished
        bool const keep = P == 1 && ld(g.keep, p) != 0;
        FT ku[NB2], kl[NB2];
#pragma unroll
        for (int q = 0; q < NB2; ++q) ku[q] = kl[q] = FT(0);
        auto put_u = [&](int k, FT const (&a)[NB2]) {
            if (keep) {
#pragma unroll
                for (int q = 0; q < NB2; ++q) ku[q] = a[q];
            } else {
                lu.store(k, a);
            }
        };
        auto put_l = [&](int e, FT const (&a)[NB2]) {
            if (keep) {
#pragma unroll
                for (int q = 0; q < NB2; ++q) kl[q] = a[q];
            } else {
                lu.store(e, a);
            }
        };
        auto zero_u = [&](int k) {
            if (keep) {
#pragma unroll
                for (int q = 0; q < NB2; ++q) ku[q] = FT(0);
            } else {
                lu.store_zero(k);
            }
        };
        auto zero_l = [&](int e) {
            if (keep) {
#pragma unroll
                for (int q = 0; q < NB2; ++q) kl[q] = FT(0);
            } else {
                lu.store_zero(e);
            }
        };
        double D[NB2];
        double rhs[NB];
#pragma unroll
        for (int k = 0; k < NB2; ++k) D[k] = 0.0;
#pragma unroll
        for (int k = 0; k < NB; ++k) rhs[k] = 0.0;

        if constexpr (MODE == MODE_RAW || MODE == MODE_RAWP) {
#pragma unroll
            for (int k = 0; k < NB2; ++k) D[k] = g.raw_a[(size_t(dg) * NB2 + k) * sp + s];
#pragma unroll
            for (int k = 0; k < NB; ++k) rhs[k] = g.raw_b[(size_t(p) * NB + k) * sp + s];
            for (int k = dg + es; k < re; k += es) {
                FT a[NB2];
#pragma unroll
                for (int q = 0; q < NB2; ++q) a[q] = FT(g.raw_a[(size_t(k) * NB2 + q) * sp + s]);
                put_u(k, a);
                int const e = ld(g.transpose, k);
#pragma unroll
                for (int q = 0; q < NB2; ++q) a[q] = FT(g.raw_a[(size_t(e) * NB2 + q) * sp + s]);
                put_l(e, a);
            }
        } else if constexpr (MODE == MODE_INIT) {
            // admittance matrix in the real domain plus load admittances and sources
            for (int k = dg; k < re; k += es) {
                FT a[NB2];
                if (ld(g.fill, k) || dead_p) {
#pragma unroll
                    for (int q = 0; q < NB2; ++q) a[q] = FT(0);
                } else {
                    double yk[PP * 2];
                    load_y<P, UNI>(g, k, s, yk);
                    linear_block<P, FT>(yk, a);
                }
                if (k == dg) {
#pragma unroll
                    for (int q = 0; q < NB2; ++q) D[q] = double(a[q]);
                } else {
                    put_u(k, a);
                    int const e = ld(g.transpose, k);
                    if (ld(g.fill, e) || dead_p) {
                        zero_l(e);
                    } else {
                        double ye[PP * 2];
                        load_y<P, UNI>(g, e, s, ye);
                        linear_block<P, FT>(ye, a);
                        put_l(e, a);
                    }
                }
            }
            if (dead_p) {
#pragma unroll
                for (int q = 0; q < NB2; ++q) D[q] = (q % (NB + 1) == 0) ? 1.0 : 0.0;
            }
            for (int l = ld(g.load_ptr, p); l < (dead_p ? ld(g.load_ptr, p) : ld(g.load_ptr, p + 1)); ++l) {
#pragma unroll
                for (int ph = 0; ph < P; ++ph) {
                    double const sr = g.s_inj[((size_t(l) * P + ph) * 2) * sp + s];
                    double const si = g.s_inj[((size_t(l) * P + ph) * 2 + 1) * sp + s];
                    // y_load = -conj(s)
                    double const yr = -sr, yi = si;
                    D[(ph + P) * NB + ph] += -yi;
                    D[ph * NB + ph] += yr;
                    D[(ph + P) * NB + ph + P] += yr;
                    D[ph * NB + ph + P] += yi;
                }
            }
            for (int q = ld(g.src_ptr, p); q < (dead_p ? ld(g.src_ptr, p) : ld(g.src_ptr, p + 1)); ++q) {
                cplx const y1 = {g.y_src[(size_t(q) * 4) * sp + s], g.y_src[(size_t(q) * 4 + 1) * sp + s]};
                cplx const y0 = {g.y_src[(size_t(q) * 4 + 2) * sp + s], g.y_src[(size_t(q) * 4 + 3) * sp + s]};
                cplx const ur = {g.u_src[(size_t(q) * 2) * sp + s], g.u_src[(size_t(q) * 2 + 1) * sp + s]};
                double ys[PP * 2];
                source_y<P>(y1, y0, ys);
                cplx ub[P];
                balanced<P>(ur, ub);
#pragma unroll
                for (int r = 0; r < P; ++r) {
                    cplx acc = {0.0, 0.0};
#pragma unroll
                    for (int c = 0; c < P; ++c) {
                        double const gg = ys[(r * P + c) * 2], bb = ys[(r * P + c) * 2 + 1];
                        D[(c + P) * NB + r] -= bb;
                        D[c * NB + r] += gg;
                        D[(c + P) * NB + r + P] += gg;
                        D[c * NB + r + P] += bb;
                        acc = cadd(acc, cmul({gg, bb}, ub[c]));
                    }
                    rhs[r] += acc.re;
                    rhs[r + P] += acc.im;
                }
            }
            if constexpr (UNI && P == 1) {
                if (g.xbus && ld(g.xbus, p) && !dead_p) {
                    cplx const none[1] = {{0.0, 0.0}};
                    double unused_p = 0.0, unused_q = 0.0;
                    extra_entries<MODE_INIT>(g, p, s, none, unused_p, unused_q);
                }
            }
        } else if constexpr (MODE == MODE_ICF || MODE == MODE_RESID) {
            // iterative current in residual form: r_p = I_p(u) - sum_j Y_pj u_j - y_src u_p, with I_p the
            // injected currents of the loads (by type) and sources at the present voltages
            // (iterative_current_pf_solver.hpp prepare_matrix_and_rhs); ICF also assembles Y + y_src
            cplx up[P], r[P];
            load_u<P>(g, p, s, up);
#pragma unroll
            for (int ph = 0; ph < P; ++ph) r[ph] = {0.0, 0.0};
            if (dead_p) {
                if constexpr (MODE == MODE_ICF) {
                    for (int k = dg + es; k < re; k += es) {
                        zero_u(k);
                        zero_l(ld(g.transpose, k));
                    }
#pragma unroll
                    for (int q = 0; q < NB2; ++q) D[q] = (q % (NB + 1) == 0) ? 1.0 : 0.0;
                }
            } else {
                for (int k = rb; k < re; k += es) {
                    int const j = ld(g.col, k);
                    if (ld(g.fill, k)) {
                        if constexpr (MODE == MODE_ICF) {
                            if (j > p) {
                                zero_u(k);
                                zero_l(ld(g.transpose, k));
                            }
                        }
                        continue;
                    }
                    double yk[PP * 2];
                    load_y<P, UNI>(g, k, s, yk);
                    cplx uj[P];
                    if (j == p) {
#pragma unroll
                        for (int ph = 0; ph < P; ++ph) uj[ph] = up[ph];
                    } else {
                        load_u<P>(g, j, s, uj);
                    }
#pragma unroll
                    for (int rr = 0; rr < P; ++rr)
#pragma unroll
                        for (int c = 0; c < P; ++c)
                            r[rr] = csub(r[rr], cmul({yk[(rr * P + c) * 2], yk[(rr * P + c) * 2 + 1]}, uj[c]));
                    if constexpr (MODE == MODE_ICF) {
                        FT a[NB2];
                        linear_block<P, FT>(yk, a);
                        if (j == p) {
#pragma unroll
                            for (int q = 0; q < NB2; ++q) D[q] = double(a[q]);
                        } else if (j > p) {
                            put_u(k, a);
                            int const e = ld(g.transpose, k);
                            if (ld(g.fill, e)) {
                                zero_l(e);
                            } else {
                                double ye[PP * 2];
                                load_y<P, UNI>(g, e, s, ye);
                                linear_block<P, FT>(ye, a);
                                put_l(e, a);
                            }
                        }
                    }
                }
                for (int l = ld(g.load_ptr, p); l < ld(g.load_ptr, p + 1); ++l) {
                    int const type = ld(g.load_type, l);
#pragma unroll
                    for (int ph = 0; ph < P; ++ph) {
                        cplx const sv = {g.s_inj[((size_t(l) * P + ph) * 2) * sp + s],
                                         g.s_inj[((size_t(l) * P + ph) * 2 + 1) * sp + s]};
                        cplx const u = up[ph];
                        cplx i;
                        if (type == 0) {
                            i = cconj(cdiv(sv, u));
                        } else if (type == 1) {
                            i = cmul(cconj(sv), u);
                        } else {
                            double const a = sqrt(u.re * u.re + u.im * u.im);
                            i = cconj(cdiv({sv.re * a, sv.im * a}, u));
                        }
                        r[ph] = cadd(r[ph], i);
                    }
                }
                for (int q = ld(g.src_ptr, p); q < ld(g.src_ptr, p + 1); ++q) {
                    cplx const y1 = {g.y_src[(size_t(q) * 4) * sp + s], g.y_src[(size_t(q) * 4 + 1) * sp + s]};
                    cplx const y0 = {g.y_src[(size_t(q) * 4 + 2) * sp + s], g.y_src[(size_t(q) * 4 + 3) * sp + s]};
                    cplx const ur = {g.u_src[(size_t(q) * 2) * sp + s], g.u_src[(size_t(q) * 2 + 1) * sp + s]};
                    double ys[PP * 2];
                    source_y<P>(y1, y0, ys);
                    cplx ub[P];
                    balanced<P>(ur, ub);
#pragma unroll
                    for (int rr = 0; rr < P; ++rr)
#pragma unroll
                        for (int c = 0; c < P; ++c) {
                            double const gg = ys[(rr * P + c) * 2], bb = ys[(rr * P + c) * 2 + 1];
                            r[rr] = cadd(r[rr], cmul({gg, bb}, csub(ub[c], up[c])));
                            if constexpr (MODE == MODE_ICF) {
                                D[(c + P) * NB + rr] -= bb;
                                D[c * NB + rr] += gg;
                                D[(c + P) * NB + rr + P] += gg;
                                D[c * NB + rr + P] += bb;
                            }
                        }
                }
            }
#pragma unroll
            for (int ph = 0; ph < P; ++ph) {
                rhs[ph] = r[ph].re;
                rhs[ph + P] = r[ph].im;
            }
            if constexpr (MODE == MODE_RESID) {
                // stored factors: y_p = L_pp^-1 P (r_p - sum_k L_pk y_k)
                for (int k = rb; k < dg; k += es) {
                    int const j = ld(g.col, k);
                    FT L[NB2];
                    lu.load(k, L);
                    double yj[NB];
#pragma unroll
                    for (int rr = 0; rr < NB; ++rr) yj[rr] = g.y[(size_t(j) * NB + rr) * sp + s];
#pragma unroll
                    for (int rr = 0; rr < NB; ++rr) {
                        double acc = 0.0;
#pragma unroll
                        for (int c = 0; c < NB; ++c) acc += double(L[c * NB + rr]) * yj[c];
                        rhs[rr] -= acc;
                    }
                }
                FT Dst[NB2];
                lu.load(dg, Dst);
                uint32_t const code = g.perm[size_t(p) * sp + s];
                apply_p_vec<NB>(rhs, code);
                lower_vec<NB, FT>(Dst, rhs);
#pragma unroll
                for (int rr = 0; rr < NB; ++rr) g.y[(size_t(p) * NB + rr) * sp + s] = rhs[rr];
                continue;
            }
        } else if (dead_p) {
            for (int k = dg + es; k < re; k += es) {
                zero_u(k);
                zero_l(ld(g.transpose, k));
            }
#pragma unroll
            for (int q = 0; q < NB2; ++q) D[q] = (q % (NB + 1) == 0) ? 1.0 : 0.0;
        } else {
            // Newton-Raphson Jacobian row p, column p and power mismatch
            cplx up[P];
            bool const from_child = kFwd && fw_uok;  // the previous row is a child handing over
            if (from_child) {
                up[0] = {fw_u[lane][2], fw_u[lane][3]};
            } else {
                load_u<P>(g, p, s, up);
            }
            double pcal[P], qcal[P];
#pragma unroll
            for (int ph = 0; ph < P; ++ph) pcal[ph] = qcal[ph] = 0.0;
            for (int k = rb; k < re; k += es) {
                int const j = ld(g.col, k);
                if (ld(g.fill, k)) {
                    if (j > p) {
                        zero_u(k);
                        zero_l(ld(g.transpose, k));
                    }
                    continue;
                }
                cplx uj[P];
                if (j == p) {
#pragma unroll
                    for (int ph = 0; ph < P; ++ph) uj[ph] = up[ph];
                } else if (from_child && j == ld(g.col, fw_le)) {
                    uj[0] = {fw_u[lane][0], fw_u[lane][1]};
                } else {
                    load_u<P>(g, j, s, uj);
                }
                if constexpr (kFwd) {
                    if (j > p && keep) {
                        // the row's only U entry: the parent's voltage, kept for the parent
                        fw_u[lane][2] = uj[0].re;
                        fw_u[lane][3] = uj[0].im;
                    }
                }
                double pr[P], qr[P];
                double yk[PP * 2];
                load_y<P, UNI>(g, k, s, yk);
                if (j == p) {
                    hnml<P, double>(yk, up, uj, D, pr, qr);
                } else {
                    FT a[NB2];
                    hnml<P, FT>(yk, up, uj, a, pr, qr);
                    if (j > p) put_u(k, a);
                }
#pragma unroll
                for (int ph = 0; ph < P; ++ph) {
                    pcal[ph] += pr[ph];
                    qcal[ph] += qr[ph];
                }
                if (j > p) {
                    int const e = ld(g.transpose, k);
                    if (ld(g.fill, e)) {
                        zero_l(e);
                    } else {
                        FT a[NB2];
                        double pr2[P], qr2[P];
                        double ye[PP * 2];
                        load_y<P, UNI>(g, e, s, ye);
                        hnml<P, FT>(ye, uj, up, a, pr2, qr2);
                        put_l(e, a);
                    }
                }
            }
            if constexpr (UNI && P == 1) {
                if (g.xbus && ld(g.xbus, p)) extra_entries<MODE_NR>(g, p, s, up, pcal[0], qcal[0]);
            }
            if constexpr (kFwd) {
                if (keep) {
                    fw_u[lane][0] = up[0].re;
                    fw_u[lane][1] = up[0].im;
                }
            }
            nr_bus_terms<P>(g, p, s, up, pcal, qcal, D, rhs);
        }

        // Schur updates owned by pivot p, pulled from the already factorised descendants. The diagonal
        // ones come one per L entry (p, k) of the row, k ascending, and also carry the forward
        // substitution b_p -= L_pk y_k
        FT Df[NB2];
#pragma unroll
        for (int k = 0; k < NB2; ++k) Df[k] = FT(D[k]);
        // FP32 factors of a Newton-Raphson sweep (plain and packed groups): the perturbation threshold from the
        // row's scale, the largest entry of its assembled diagonal block (the Jacobian's rows are close to
        // diagonally dominant, so this is about the row's norm)
        // (kept in LDS through the Schur updates, which hold every register they can get)
        if constexpr (kPerturb) {
            double scale = 0.0;
#pragma unroll
            for (int q = 0; q < NB2; ++q) scale = fabs(D[q]) > scale ? fabs(D[q]) : scale;
            thr_s[lane] = FT(kPivotPerturbation * scale);
        }
        for (int q = ld(g.schur_ptr, p); q < ld(g.schur_ptr, p + 1); ++q) {
            int const tgt = ld(g.schur, 3 * q), le = ld(g.schur, 3 * q + 1), ue = ld(g.schur, 3 * q + 2);
            FT L[NB2], U[NB2];
            bool const fwd = kFwd && le == fw_le;
            if (fwd) {
#pragma unroll
                for (int k = 0; k < NB2; ++k) {
                    L[k] = fw_blk[lane][k];
                    U[k] = fw_blk[lane][NB2 + k];
                }
            } else {
                lu.load(le, L);
                lu.load(ue, U);
            }
            if (tgt == dg) {
#pragma unroll
                for (int c = 0; c < NB; ++c)
#pragma unroll
                    for (int r = 0; r < NB; ++r) {
                        FT acc = FT(0);
#pragma unroll
                        for (int k = 0; k < NB; ++k) acc += L[k * NB + r] * U[c * NB + k];
                        Df[c * NB + r] -= acc;
                    }
                double yj[NB];
                if (fwd) {
#pragma unroll
                    for (int r = 0; r < NB; ++r) yj[r] = fw_y[lane][r];
                } else {
                    int const j = ld(g.col, le);
#pragma unroll
                    for (int r = 0; r < NB; ++r) yj[r] = g.y[(size_t(j) * NB + r) * sp + s];
                }
#pragma unroll
                for (int r = 0; r < NB; ++r) {
                    double acc = 0.0;
#pragma unroll
                    for (int c = 0; c < NB; ++c) acc += double(L[c * NB + r]) * yj[c];
                    rhs[r] -= acc;
                }
            } else {
                FT A[NB2];
                lu.load(tgt, A);
#pragma unroll
                for (int c = 0; c < NB; ++c)
#pragma unroll
                    for (int r = 0; r < NB; ++r) {
                        FT acc = FT(0);
#pragma unroll
                        for (int k = 0; k < NB; ++k) acc += L[k * NB + r] * U[c * NB + k];
                        A[c * NB + r] -= acc;
                    }
                lu.store(tgt, A);
            }
        }

        uint32_t code;
        // a pivot the FP32 factor cannot carry is perturbed (the step is then refined in FP64, see
        // refine_residual_kernel); without perturbation the pivot test is the reference's
        bool perturbed = false;
        // FP64 raw systems (the state-estimation gain matrix) perturb below a threshold per pair when one is
        // given (sparse_lu_solver.hpp:39-48, 352)
        constexpr bool kPerturbRaw = (MODE == MODE_RAW || MODE == MODE_RAWP) && sizeof(FT) == 8;
        FT raw_thr = FT(0);
        if constexpr (kPerturbRaw) raw_thr = g.pthr ? FT(g.pthr[size_t(mm) * sp + s]) : FT(0);
        bool ok;
        if constexpr (MODE == MODE_RAWP)
            ok = factor_block<NB, FT>(Df, code, raw_thr, perturbed, raw_pert);
        else
            ok = factor_block<NB, FT>(Df, code, kPerturb ? thr_s[lane] : (kPerturbRaw ? raw_thr : FT(0)), perturbed);
        if constexpr (kPerturb) {
            // 2: a pivot was perturbed, the refinement must reach its bound
            if (perturbed) atomicMax(&g.pert[size_t(mm) * sp + s], 2);
        }
        if constexpr (kPerturbRaw) {
            if (perturbed && g.pert) atomicMax(&g.pert[size_t(mm) * sp + s], 1);
        }
        if constexpr (MODE == MODE_RAWP) raw_pert = raw_pert || perturbed;
        if (!ok) {
            if constexpr (MODE != MODE_NRA) {
                g.done[size_t(mm) * sp + s] = FAILED;
                g.err[size_t(mm) * sp + s] = ERR_SPARSE;
            }
            return;
        }
        lu.store(dg, Df);
        g.perm[size_t(p) * sp + s] = code;
        // a kept row whose parent is the task's next pivot hands its results over through LDS
        int const t_end = ld(g.task_ptr, task + 1);
        bool const hand = kFwd && keep && ti + 1 < t_end && ld(g.task_pivot, ti + 1) == ld(g.col, dg + es);
        // U row: U_pj = L^-1 P A_pj ; L column: L_ip = A_ip Q U^-1
        for (int k = dg + es; k < re; k += es) {
            FT A[NB2];
            if (keep) {
#pragma unroll
                for (int q = 0; q < NB2; ++q) A[q] = ku[q];
            } else {
                lu.load(k, A);
            }
            apply_p_rows<NB, FT>(A, code);
            lower_left<NB, FT>(Df, A);
            lu.store(k, A);
            if (kFwd && keep) {
#pragma unroll
                for (int q = 0; q < NB2; ++q) fw_blk[lane][NB2 + q] = A[q];
            }
            int const e = ld(g.transpose, k);
            if (keep) {
#pragma unroll
                for (int q = 0; q < NB2; ++q) A[q] = kl[q];
            } else {
                lu.load(e, A);
            }
            apply_q_cols<NB, FT>(A, code);
            upper_right<NB, FT>(Df, A);
            if (!(kSkipL && hand)) lu.store(e, A);
            if (kFwd && keep) {
#pragma unroll
                for (int q = 0; q < NB2; ++q) fw_blk[lane][q] = A[q];
            }
        }
        // forward substitution: y_p = L_pp^-1 P (b_p - sum_k L_pk y_k)
        apply_p_vec<NB>(rhs, code);
        lower_vec<NB, FT>(Df, rhs);
#pragma unroll
        for (int r = 0; r < NB; ++r) g.y[(size_t(p) * NB + r) * sp + s] = rhs[r];
        if constexpr (kFwd) {
            fw_le = -1;
            fw_uok = false;
            if (hand) {
                int const k1 = dg + es;  // the row's only U entry
                fw_le = ld(g.transpose, k1);
#pragma unroll
                for (int r = 0; r < NB; ++r) fw_y[lane][r] = rhs[r];
                // voltages of the row and its parent (read by this row's U entry)
                fw_uok = MODE == MODE_NR && !dead_p && !ld(g.fill, k1);
            }
        }
    }
    if constexpr (kPerturb) {
        // 1: refined on request only
        if (g.refine_all) atomicMax(&g.pert[size_t(mm) * sp + s], 1);
    }
}

// ---------------------------------------------------------------------------------------------------
// backward sweep: x_p = Q U_pp^-1 (y_p - sum_j U_pj x_j), then the state update of bus p
// ---------------------------------------------------------------------------------------------------
template <int P, class FT, int MODE, int KIND>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(backward_waves<P, MODE, KIND>()))) void backward_kernel(DevGroup const* __restrict__ groups,
                                                         WorkItem const* __restrict__ items) {
    constexpr int NB = 2 * P;
    constexpr int NB2 = NB * NB;
    WorkItem const wi = ld(items, blockIdx.x);
    DevGroup const g = ld(groups, wi.group);
    int task, s;
    if (!lane_work<KIND>(g, wi, task, s)) return;
    size_t const sp = g.S_pad;
    int const mm = ld(g.task_mm, task);
    if (g.done[size_t(mm) * sp + s] != (MODE == MODE_REFINE ? REFINE : ACTIVE)) return;
    // Newton-Raphson: a step from perturbed pivots is not applied here; it waits for its refinement
    bool const held = MODE == MODE_NR && sizeof(FT) == 4 && KIND != KIND_UNION && g.pert &&
                      g.pert[size_t(mm) * sp + s] != 0;
    Blocks<NB, FT> const lu{static_cast<FT*>(g.lu), sp, s};
    double dev = 0.0;
    // x of the pivot solved last, handed to the next pivot through LDS (its parent's value when the
    // next pivot is that pivot's first child)
    __shared__ double fx[kWave][NB];
    int const lane = int(threadIdx.x);
    int fx_bus = -1;
    for (int ti = ld(g.task_ptr, task + 1) - 1; ti >= ld(g.task_ptr, task); --ti) {
        int const p = ld(g.task_pivot, ti);
        int const re = ld(g.row_end, p), dg = ld(g.diag, p);
        int const es = g.row_stride ? ld(g.row_stride, p) : 1;
        double v[NB];
#pragma unroll
        for (int r = 0; r < NB; ++r) v[r] = g.y[(size_t(p) * NB + r) * sp + s];
        for (int k = dg + es; k < re; k += es) {
            int const j = ld(g.col, k);
            FT U[NB2];
            lu.load(k, U);
            double xj[NB];
            if (j == fx_bus) {
#pragma unroll
                for (int r = 0; r < NB; ++r) xj[r] = fx[lane][r];
            } else {
#pragma unroll
                for (int r = 0; r < NB; ++r) xj[r] = g.y[(size_t(j) * NB + r) * sp + s];
            }
#pragma unroll
            for (int r = 0; r < NB; ++r) {
                double acc = 0.0;
#pragma unroll
                for (int c = 0; c < NB; ++c) acc += double(U[c * NB + r]) * xj[c];
                v[r] -= acc;
            }
        }
        FT Df[NB2];
        lu.load(dg, Df);
        upper_vec<NB, FT>(Df, v);
        apply_q_vec<NB>(v, g.perm[size_t(p) * sp + s]);
#pragma unroll
        for (int r = 0; r < NB; ++r) g.y[(size_t(p) * NB + r) * sp + s] = v[r];
        fx_bus = p;
#pragma unroll
        for (int r = 0; r < NB; ++r) fx[lane][r] = v[r];
        if constexpr (MODE == MODE_INIT) {
#pragma unroll
            for (int ph = 0; ph < P; ++ph) {
                double const ur = v[ph], ui = v[ph + P];
                size_t const b = (size_t(p) * P + ph) * 2;
                g.u[b * sp + s] = ur;
                g.u[(b + 1) * sp + s] = ui;
                g.xv[b * sp + s] = atan2(ui, ur);
                g.xv[(b + 1) * sp + s] = sqrt(ur * ur + ui * ui);
            }
        } else if constexpr (MODE == MODE_ICF || MODE == MODE_RESID) {
            // u += x; the deviation is |x| (iterative_current_pf_solver.hpp iterate_unknown)
            double dbus = 0.0;
#pragma unroll
            for (int ph = 0; ph < P; ++ph) {
                size_t const b = (size_t(p) * P + ph) * 2;
                g.u[b * sp + s] += v[ph];
                g.u[(b + 1) * sp + s] += v[ph + P];
                double const d = sqrt(v[ph] * v[ph] + v[ph + P] * v[ph + P]);
                dbus = d > dbus ? d : dbus;
            }
            dev = dbus > dev ? dbus : dev;
        } else if constexpr (MODE == MODE_REFINE) {
#pragma unroll
            for (int r = 0; r < NB; ++r) g.xacc[(size_t(p) * NB + r) * sp + s] += v[r];
        } else if constexpr (MODE == MODE_NR) {
            if (held) continue;
            double dbus = 0.0;
#pragma unroll
            for (int ph = 0; ph < P; ++ph) {
                size_t const b = (size_t(p) * P + ph) * 2;
                double const th0 = g.xv[b * sp + s], vm0 = g.xv[(b + 1) * sp + s];
                double th = th0 + v[ph];
                double vm = vm0;
                vm += vm * v[ph + P];
                double sn, cs, sn0, cs0;
                sincos(th, &sn, &cs);
                // the previous voltage from the previous polar state, as the previous update computed it
                // (no read of the stored voltage; after the linear initial guess it differs from the stored
                // one by rounding only)
                sincos(th0, &sn0, &cs0);
                double const nr = vm * cs, ni = vm * sn;
                double const dr = nr - vm0 * cs0, di = ni - vm0 * sn0;
                double const d = sqrt(dr * dr + di * di);
                dbus = ph == 0 ? d : (d > dbus ? d : dbus);
                g.xv[b * sp + s] = th;
                g.xv[(b + 1) * sp + s] = vm;
                g.u[b * sp + s] = nr;
                g.u[(b + 1) * sp + s] = ni;
            }
            dev = dbus > dev ? dbus : dev;
        }
    }
    if constexpr (MODE == MODE_NR || MODE == MODE_ICF || MODE == MODE_RESID) {
        if (!held)
            atomicMax(&g.devmax[size_t(mm) * sp + s], static_cast<unsigned long long>(__double_as_longlong(dev)));
    }
}

// ---------------------------------------------------------------------------------------------------
// low-rank correction of a union lane with extra entries. With W the block columns of the lane's m
// distinct extra buses and M (2m x 2m) the extra blocks, the lane's matrix is J = J' + W M W^T, J' the
// factorised one. Writing the forward sweep as y = F r and the backward sweep as x = B y (so that
// J'^-1 = B F), Woodbury gives x = B (y - F W z) with z = (I + M Z)^-1 M W^T B y and Z = W^T B F W.
// On a radial structure F W and W^T B live on the chains from the buses to the root: S_k, the forward
// response at chain step k to a unit block at the chain's bus, and T_k, the response of x at the bus
// to y at step k:
//   S_0 = F_c, S_k = F_k (-L_{k,k-1} S_{k-1});  T_0 = B_c, T_k = -T_{k-1} U_{k-1,k} B_k
// (F_k: row permutation and unit-lower solve of the pivot block; B_k: upper solve and column
// permutation). Then (W^T B y)_c = sum_k T^c_k y_k, Z_cd = sum over the common steps of T^c S^d, and
// the correction y -= F W z runs along the chains before the backward sweep.
// One block (one wave) per lane.
// ---------------------------------------------------------------------------------------------------
__device__ __forceinline__ void mul22(double const (&a)[4], double const (&b)[4], double (&c)[4]) {
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 2; ++r) c[j * 2 + r] = a[r] * b[j * 2] + a[2 + r] * b[j * 2 + 1];
}

template <class FT>
__global__ __launch_bounds__(kWave) void lowrank_kernel(DevGroup const* __restrict__ groups,
                                                        ExtraWork const* __restrict__ work) {
    constexpr int K2 = 2 * kExtraMax;
    ExtraWork const w = ld(work, blockIdx.x);
    DevGroup const g = ld(groups, w.group);
    int const s = w.lane;
    size_t const sp = g.S_pad;
    int const m = ld(g.xm, s), nx = ld(g.xcnt, s);
    int const* xb = g.xb + size_t(s) * kExtraMax;
    int const mm = ld(g.bus_mm, ld(xb, 0));
    if (g.done[size_t(mm) * sp + s] != ACTIVE) return;
    Blocks<2, FT> const lu{static_cast<FT*>(g.lu), sp, s};
    double* const st = g.xst + size_t(w.slot) * kExtraMax * size_t(g.ch_max) * 8;
    int const t = int(threadIdx.x);
    __shared__ double Z[K2][K2], Mx[K2][K2], Km[K2][K2 + 1], w0[K2], z[K2];
    __shared__ int fail;

    // chains: S and T of every step
    if (t < m) {
        int const c = ld(xb, t);
        int const off = ld(g.ch_off, c), len = ld(g.ch_len, c);
        double* out = st + size_t(t) * g.ch_max * 8;
        double S[4] = {1.0, 0.0, 0.0, 1.0}, T[4] = {0.0, 0.0, 0.0, 0.0};
        for (int k = 0; k < len; ++k) {
            int const i = ld(g.ch_node, off + k);
            FT Df[4];
            lu.load(ld(g.diag, i), Df);
            uint32_t const code = g.perm[size_t(i) * sp + s];
            double A[4] = {1.0, 0.0, 0.0, 1.0}, R[4] = {1.0, 0.0, 0.0, 1.0};
            if (k > 0) {
                int const ue = ld(g.ch_ue, off + k);
                FT Lf[4], Uf[4];
                lu.load(ld(g.transpose, ue), Lf);
                lu.load(ue, Uf);
                double const Lb[4] = {double(Lf[0]), double(Lf[1]), double(Lf[2]), double(Lf[3])};
                double const Ub[4] = {double(Uf[0]), double(Uf[1]), double(Uf[2]), double(Uf[3])};
                mul22(Lb, S, A);
#pragma unroll
                for (int q = 0; q < 4; ++q) A[q] = -A[q];
                mul22(T, Ub, R);
#pragma unroll
                for (int q = 0; q < 4; ++q) R[q] = -R[q];
            }
            double B[4];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                double v[2] = {A[j * 2], A[j * 2 + 1]};
                apply_p_vec<2>(v, code);
                lower_vec<2, FT>(Df, v);
                S[j * 2] = v[0];
                S[j * 2 + 1] = v[1];
                double e[2] = {j == 0 ? 1.0 : 0.0, j == 1 ? 1.0 : 0.0};
                upper_vec<2, FT>(Df, e);
                apply_q_vec<2>(e, code);
                B[j * 2] = e[0];
                B[j * 2 + 1] = e[1];
            }
            mul22(R, B, T);  // k = 0: R = I, T = B_c
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                out[size_t(k) * 8 + q] = S[q];
                out[size_t(k) * 8 + 4 + q] = T[q];
            }
        }
    }
    for (int q = t; q < K2 * K2; q += kWave) Mx[q / K2][q % K2] = 0.0;
    __syncthreads();

    // Z over the common chain steps (one thread per bus pair), W^T B y, M
    if (t < m * m) {
        int const c = t / m, d = t % m;
        int const bc = ld(xb, c), bd = ld(xb, d);
        int const oc = ld(g.ch_off, bc), od = ld(g.ch_off, bd);
        int kc = ld(g.ch_len, bc) - 1, kd = ld(g.ch_len, bd) - 1;
        double const* sc = st + size_t(c) * g.ch_max * 8;
        double const* sd = st + size_t(d) * g.ch_max * 8;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        while (kc >= 0 && kd >= 0 && ld(g.ch_node, oc + kc) == ld(g.ch_node, od + kd)) {
            double const Tc[4] = {sc[kc * 8 + 4], sc[kc * 8 + 5], sc[kc * 8 + 6], sc[kc * 8 + 7]};
            double const Sd[4] = {sd[kd * 8], sd[kd * 8 + 1], sd[kd * 8 + 2], sd[kd * 8 + 3]};
            double prod[4];
            mul22(Tc, Sd, prod);
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] += prod[q];
            --kc;
            --kd;
        }
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 2; ++r) Z[2 * c + r][2 * d + j] = acc[j * 2 + r];
    }
    if (t < m) {
        int const c = ld(xb, t);
        int const off = ld(g.ch_off, c), len = ld(g.ch_len, c);
        double const* sc = st + size_t(t) * g.ch_max * 8;
        double a0 = 0.0, a1 = 0.0;
        for (int k = 0; k < len; ++k) {
            int const i = ld(g.ch_node, off + k);
            double const y0 = g.y[(size_t(i) * 2) * sp + s], y1 = g.y[(size_t(i) * 2 + 1) * sp + s];
            a0 += sc[k * 8 + 4] * y0 + sc[k * 8 + 6] * y1;
            a1 += sc[k * 8 + 5] * y0 + sc[k * 8 + 7] * y1;
        }
        w0[2 * t] = a0;
        w0[2 * t + 1] = a1;
    }
    if (t < nx) {
        size_t const q = size_t(s) * kExtraMax + t;
        int const rb = ld(g.xrow, q), cb = ld(g.xcol, q);
        int ri = 0, ci = 0;
        for (int k = 0; k < m; ++k) {
            int const b = ld(xb, k);
            ri = b == rb ? k : ri;
            ci = b == cb ? k : ci;
        }
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 2; ++r) Mx[2 * ri + r][2 * ci + j] = g.xblk[q * 4 + j * 2 + r];
    }
    __syncthreads();

    // (I + M Z) z = M w0, Gaussian elimination with partial pivoting
    if (t == 0) {
        int const n2 = 2 * m;
        for (int r = 0; r < n2; ++r) {
            double rhs = 0.0;
            for (int k = 0; k < n2; ++k) rhs += Mx[r][k] * w0[k];
            Km[r][n2] = rhs;
            for (int c = 0; c < n2; ++c) {
                double acc = r == c ? 1.0 : 0.0;
                for (int k = 0; k < n2; ++k) acc += Mx[r][k] * Z[k][c];
                Km[r][c] = acc;
            }
        }
        int bad = 0;
        for (int c = 0; c < n2; ++c) {
            int pr = c;
            for (int r = c + 1; r < n2; ++r)
                if (fabs(Km[r][c]) > fabs(Km[pr][c])) pr = r;
            if (!(fabs(Km[pr][c]) > 0.0) || !isfinite(Km[pr][c])) {
                bad = 1;
                break;
            }
            if (pr != c)
                for (int k = c; k <= n2; ++k) {
                    double const x = Km[c][k];
                    Km[c][k] = Km[pr][k];
                    Km[pr][k] = x;
                }
            for (int r = c + 1; r < n2; ++r) {
                double const f = Km[r][c] / Km[c][c];
                for (int k = c; k <= n2; ++k) Km[r][k] -= f * Km[c][k];
            }
        }
        if (!bad)
            for (int r = n2 - 1; r >= 0; --r) {
                double v = Km[r][n2];
                for (int k = r + 1; k < n2; ++k) v -= Km[r][k] * z[k];
                z[r] = v / Km[r][r];
            }
        fail = bad;
        if (bad) {
            g.done[size_t(mm) * sp + s] = FAILED;
            g.err[size_t(mm) * sp + s] = ERR_SPARSE;
        }
    }
    __syncthreads();
    if (fail) return;

    // y -= F W z along the chains, one bus after the other (chains share their upper steps)
    for (int d = 0; d < m; ++d) {
        int const bd = ld(xb, d);
        int const off = ld(g.ch_off, bd), len = ld(g.ch_len, bd);
        double const* sd = st + size_t(d) * g.ch_max * 8;
        double const z0 = z[2 * d], z1 = z[2 * d + 1];
        for (int k = t; k < len; k += kWave) {
            int const i = ld(g.ch_node, off + k);
            g.y[(size_t(i) * 2) * sp + s] -= sd[k * 8] * z0 + sd[k * 8 + 2] * z1;
            g.y[(size_t(i) * 2 + 1) * sp + s] -= sd[k * 8 + 1] * z0 + sd[k * 8 + 3] * z1;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------------------
// adjoint: J^T lambda = g with the factors J = L~ U~ of the sweeps (L~: P^T L_pp on the diagonal and the
// L blocks; U~: U_pp Q^T on the diagonal and the U blocks). U~^T w = g runs in the forward order:
//   w_p = U_pp^-T Q^T (g_p - sum_k U_kp^T w_k)      over the L entries (p, k) of row p;
// then L~^T lambda = w in the backward order:
//   lambda_p = P^T L_pp^-T (w_p - sum_i L_ip^T lambda_i)   over the U entries (p, i) of row p.
// Symmetric batches, plain groups (one task, 64 scenarios per wave).
// ---------------------------------------------------------------------------------------------------
template <int NB> __device__ __forceinline__ void apply_qt_vec(double (&v)[NB], uint32_t code) {
#pragma unroll
    for (int t = 0; t < NB - 1; ++t) {
        int const c = perm_col(code, t);
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            if (k <= t) continue;
            bool const sw = c == k;
            double const x = v[t], y = v[k];
            v[t] = sw ? y : x;
            v[k] = sw ? x : y;
        }
    }
}
template <int NB> __device__ __forceinline__ void apply_pt_vec(double (&v)[NB], uint32_t code) {
#pragma unroll
    for (int t = NB - 2; t >= 0; --t) {
        int const r = perm_row(code, t);
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            if (k <= t) continue;
            bool const sw = r == k;
            double const x = v[t], y = v[k];
            v[t] = sw ? y : x;
            v[k] = sw ? x : y;
        }
    }
}
// v := U^-T v, U the upper part of a factored block (lu[c * NB + r], r <= c)
template <int NB, class FT> __device__ __forceinline__ void upper_t_vec(FT const (&lu)[NB * NB], double (&v)[NB]) {
#pragma unroll
    for (int r = 0; r < NB; ++r) {
#pragma unroll
        for (int c = 0; c < r; ++c) v[r] -= double(lu[r * NB + c]) * v[c];
        v[r] /= double(lu[r * NB + r]);
    }
}
// v := L^-T v, L the unit lower part (lu[c * NB + r], r > c)
template <int NB, class FT> __device__ __forceinline__ void lower_t_vec(FT const (&lu)[NB * NB], double (&v)[NB]) {
#pragma unroll
    for (int r = NB - 1; r >= 0; --r)
#pragma unroll
        for (int c = NB - 1; c > r; --c) v[r] -= double(lu[r * NB + c]) * v[c];
}

template <class FT, bool UT>
__global__ __launch_bounds__(kWave) void adjoint_sweep_kernel(DevGroup const* __restrict__ groups,
                                                              WorkItem const* __restrict__ items) {
    constexpr int NB = 2, NB2 = 4;
    WorkItem const wi = ld(items, blockIdx.x);
    DevGroup const g = ld(groups, wi.group);
    int const task = wi.task;
    int const s = wi.block * kWave + int(threadIdx.x);
    if (s >= g.S) return;
    size_t const sp = g.S_pad;
    int const mm = ld(g.task_mm, task);
    if (g.done[size_t(mm) * sp + s] != CONVERGED) return;
    Blocks<NB, FT> const lu{static_cast<FT*>(g.lu), sp, s};
    int const t0 = ld(g.task_ptr, task), t1 = ld(g.task_ptr, task + 1);
    for (int q = 0; q < t1 - t0; ++q) {
        int const ti = UT ? t0 + q : t1 - 1 - q;
        int const p = ld(g.task_pivot, ti);
        int const rb = ld(g.row_ptr, p), re = ld(g.row_end, p), dg = ld(g.diag, p);
        int const es = g.row_stride ? ld(g.row_stride, p) : 1;
        double v[NB];
#pragma unroll
        for (int r = 0; r < NB; ++r) v[r] = g.y[(size_t(p) * NB + r) * sp + s];
        // UT: the L entries (p, j), j < p, give U_jp = the block at their transpose; LT: the U entries
        // (p, j), j > p, give L_jp
        int const k0 = UT ? rb : dg + es, k1 = UT ? dg : re;
        for (int k = k0; k < k1; k += es) {
            int const j = ld(g.col, k);
            FT B[NB2];
            lu.load(ld(g.transpose, k), B);
            double xj[NB];
#pragma unroll
            for (int r = 0; r < NB; ++r) xj[r] = g.y[(size_t(j) * NB + r) * sp + s];
#pragma unroll
            for (int r = 0; r < NB; ++r) {
                double acc = 0.0;
#pragma unroll
                for (int c = 0; c < NB; ++c) acc += double(B[r * NB + c]) * xj[c];
                v[r] -= acc;
            }
        }
        FT Df[NB2];
        lu.load(dg, Df);
        uint32_t const code = g.perm[size_t(p) * sp + s];
        if constexpr (UT) {
            apply_qt_vec<NB>(v, code);
            upper_t_vec<NB, FT>(Df, v);
        } else {
            lower_t_vec<NB, FT>(Df, v);
            apply_pt_vec<NB>(v, code);
        }
#pragma unroll
        for (int r = 0; r < NB; ++r) g.y[(size_t(p) * NB + r) * sp + s] = v[r];
    }
}

// dL/dx at the converged state, x = (theta, ln V) as the Jacobian's columns: the voltage term
// 2 V (V - 1) per node of the bus, and the losses of the lines at the bus. For a line with admittances
// ff, ft, tf, tt: loss = Re(ff) V_f^2 + Re(A) + Re(B) + Re(tt) V_t^2 with A = conj(ft) u_f conj(u_t),
// B = conj(tf) u_t conj(u_f); d/dtheta_f = Im(B) - Im(A) = -d/dtheta_t, d/dlnV_f = 2 Re(ff) V_f^2 + Re(A) + Re(B)
__global__ void adjoint_rhs_kernel(DevGroup const* __restrict__ groups, DevAdjoint const* __restrict__ adjs,
                                   long long const* __restrict__ offsets, int n_groups) {
    long long const gidx = (long long)(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gidx >= ld(offsets, n_groups)) return;
    int const gr = find_group(offsets, n_groups, gidx);
    DevGroup const g = ld(groups, gr);
    DevAdjoint const a = ld(adjs, gr);
    size_t const sp = g.S_pad;
    size_t const idx = size_t(gidx - ld(offsets, gr));
    int const b = int(idx / sp), s = int(idx % sp);
    double gth = 0.0, glv = 0.0;
    int const mm = ld(g.bus_mm, b);
    if (s < g.S && mm >= 0 && g.done[size_t(mm) * sp + s] == CONVERGED) {
        auto U = [&](int bus) { return cplx{g.u[(size_t(bus) * 2) * sp + s], g.u[(size_t(bus) * 2 + 1) * sp + s]}; };
        double const V = g.xv[(size_t(b) * 2 + 1) * sp + s];
        glv = 2.0 * V * (V - 1.0) * double(ld(a.node_w, b));
        for (int q = ld(a.bus_line_ptr, b); q < ld(a.bus_line_ptr, b + 1); ++q) {
            int const mb = ld(a.bus_line, q);
            int const f = ld(a.br_from, mb), t = ld(a.br_to, mb);
            auto Y = [&](int w) { return cplx{ld(a.br_y, (mb * 4 + w) * 2), ld(a.br_y, (mb * 4 + w) * 2 + 1)}; };
            bool const at_f = f == b;
            cplx const own = Y(at_f ? 0 : 3);
            glv += 2.0 * own.re * V * V;
            if (f >= 0 && t >= 0 && f != t) {
                cplx const uf = U(f), ut = U(t);
                cplx const A = cmul(cconj(Y(1)), cmul(uf, cconj(ut)));
                cplx const B = cmul(cconj(Y(2)), cmul(ut, cconj(uf)));
                glv += A.re + B.re;
                gth += at_f ? (B.im - A.im) : (A.im - B.im);
            }
        }
    }
    g.y[(size_t(b) * 2) * sp + s] = gth;
    g.y[(size_t(b) * 2 + 1) * sp + s] = glv;
}

// gradients of lane s: loads dL/dp = lambda_P dir f(V) / 1e6 (f = 1, V^2, V for const power, impedance,
// current), dL/dq likewise with lambda_Q; lines dL/dr = dloss/dr - sum over both ends of
// (lambda_P Re dS + lambda_Q Im dS), dS_f = u_f conj(dff u_f + dft u_t), dS_t = u_t conj(dtf u_f + dtt u_t)
__global__ void adjoint_grad_kernel(DevGroup const* __restrict__ groups, DevOutput const* __restrict__ outs,
                                    DevAdjoint const* __restrict__ adjs, long long const* __restrict__ offsets,
                                    int n_groups) {
    long long const gidx = (long long)(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gidx >= ld(offsets, n_groups)) return;
    int const gr = find_group(offsets, n_groups, gidx);
    DevGroup const g = ld(groups, gr);
    DevOutput const o = ld(outs, gr);
    DevAdjoint const a = ld(adjs, gr);
    size_t const sp = g.S_pad;
    size_t const idx = size_t(gidx - ld(offsets, gr));
    int const c = int(idx / sp), s = int(idx % sp);
    if (s >= g.S || !o.ok[s]) return;
    long long const scn = ld(o.scenario_index, s);
    auto lam = [&](int bus, int r) { return g.y[(size_t(bus) * 2 + r) * sp + s]; };
    auto U = [&](int bus) { return cplx{g.u[(size_t(bus) * 2) * sp + s], g.u[(size_t(bus) * 2 + 1) * sp + s]}; };
    if (c < a.n_lg) {
        double gp = 0.0, gq = 0.0;
        int const k = ld(a.lg_math, c);
        if (k >= 0 && ld(a.load_status, size_t(c) * sp + s)) {
            int const bus = ld(a.load_bus, k);
            int const type = ld(g.load_type, k);
            double const V = g.xv[(size_t(bus) * 2 + 1) * sp + s];
            double const f = type == 0 ? 1.0 : (type == 1 ? V * V : V);
            double const dir = ld(a.lg_gen, c) ? 1.0 : -1.0;
            gp = lam(bus, 0) * dir * f / 1e6;
            gq = lam(bus, 1) * dir * f / 1e6;
        }
        a.grad_p[scn * a.n_lg + c] = gp;
        a.grad_q[scn * a.n_lg + c] = gq;
        return;
    }
    int const line = c - a.n_lg;
    double gr_ = 0.0, gx = 0.0;
    int const mb = ld(a.line_math, line);
    if (mb >= 0) {
        int const f = ld(a.br_from, mb), t = ld(a.br_to, mb);
        cplx const uf = f >= 0 ? U(f) : cplx{0.0, 0.0}, ut = t >= 0 ? U(t) : cplx{0.0, 0.0};
        for (int which = 0; which < 2; ++which) {
            double const* d = which == 0 ? a.br_dr : a.br_dx;
            auto D = [&](int w) { return cplx{ld(d, (mb * 4 + w) * 2), ld(d, (mb * 4 + w) * 2 + 1)}; };
            cplx const dsf = cmul(uf, cconj(cadd(cmul(D(0), uf), cmul(D(1), ut))));
            cplx const dst = cmul(ut, cconj(cadd(cmul(D(2), uf), cmul(D(3), ut))));
            double gsum = dsf.re + dst.re;
            if (f >= 0) gsum -= lam(f, 0) * dsf.re + lam(f, 1) * dsf.im;
            if (t >= 0) gsum -= lam(t, 0) * dst.re + lam(t, 1) * dst.im;
            (which == 0 ? gr_ : gx) = gsum;
        }
    }
    a.grad_r[scn * a.n_line + line] = gr_;
    a.grad_x[scn * a.n_line + line] = gx;
}

// ---------------------------------------------------------------------------------------------------
// load_gen updates on the device (update.hpp semantics: a null status or power keeps the base value;
// s = direction (p + j q) / base_power, load_gen.hpp:121-139; asymmetric components in a symmetric
// calculation take the mean of the phases)
// ---------------------------------------------------------------------------------------------------
__global__ void load_broadcast_kernel(double* __restrict__ s_inj, int8_t* __restrict__ lstat,
                                      double const* __restrict__ base_s, int8_t const* __restrict__ base_status,
                                      int rows_s, int n_lg, int S, int S_pad) {
    long long const i = (long long)(blockIdx.x) * blockDim.x + threadIdx.x;
    long long const n_s = (long long)(rows_s) * S_pad;
    if (i < n_s) {
        int const row = int(i / S_pad), lane = int(i % S_pad);
        s_inj[i] = lane < S ? base_s[row] : 0.0;
    } else if (i < n_s + (long long)(n_lg) * S_pad) {
        long long const j = i - n_s;
        int const l = int(j / S_pad), lane = int(j % S_pad);
        lstat[j] = lane < S ? base_status[l] : int8_t(0);
    }
}

// one tile of 64 lanes x 16 update elements per block: the update rows are read row-wise (each scenario's 16
// records in one piece), staged in LDS, and the lane-interleaved injections and statuses written lane-wise
constexpr int kScatterLanes = 64, kScatterElems = 16;
__global__ __launch_bounds__(256) void load_scatter_kernel(double* __restrict__ s_inj, int8_t* __restrict__ lstat,
                                                            int S, int S_pad, int const* __restrict__ scenario_index,
                                                            LoadUpdateSource const src, LoadTable const tab) {
    __shared__ double ps[kScatterLanes][kScatterElems][3], qs[kScatterLanes][kScatterElems][3];
    __shared__ int8_t ss[kScatterLanes][kScatterElems];
    int const lane0 = int(blockIdx.x) * kScatterLanes, e0 = int(blockIdx.y) * kScatterElems;
    int const nph = src.asym ? 3 : 1;
    // rows in: element fastest
    for (int t = int(threadIdx.x); t < kScatterLanes * kScatterElems; t += int(blockDim.x)) {
        int const le = t % kScatterElems, ll = t / kScatterElems;
        int const lane = lane0 + ll, e = e0 + le;
        if (lane >= S || e >= src.n_elem) continue;
        long long const row = (long long)(ld(scenario_index, lane)) * src.n_elem + e;
        ss[ll][le] = src.status ? *reinterpret_cast<int8_t const*>(src.status + row * src.status_stride) : int8_t(-128);
        for (int ph = 0; ph < nph; ++ph) {
            ps[ll][le][ph] = src.p ? reinterpret_cast<double const*>(src.p + row * src.p_stride)[ph] : __builtin_nan("");
            qs[ll][le][ph] = src.q ? reinterpret_cast<double const*>(src.q + row * src.q_stride)[ph] : __builtin_nan("");
        }
    }
    __syncthreads();
    // lanes out: lane fastest
    size_t const sp = S_pad;
    for (int t = int(threadIdx.x); t < kScatterLanes * kScatterElems; t += int(blockDim.x)) {
        int const ll = t % kScatterLanes, le = t / kScatterLanes;
        int const lane = lane0 + ll, e = e0 + le;
        if (lane >= S || e >= src.n_elem) continue;
        int const l = ld(src.comp, e);
        bool st = ld(tab.status, l) != 0;
        double p[3], q[3];
#pragma unroll
        for (int ph = 0; ph < 3; ++ph) {
            p[ph] = ld(tab.p, 3 * l + ph);
            q[ph] = ld(tab.q, 3 * l + ph);
        }
        int8_t const v = ss[ll][le];
        if (v != int8_t(-128)) st = v != 0;
        for (int ph = 0; ph < nph; ++ph) {
            if (!isnan(ps[ll][le][ph])) p[ph] = ps[ll][le][ph];
            if (!isnan(qs[ll][le][ph])) q[ph] = qs[ll][le][ph];
        }
        lstat[size_t(l) * sp + lane] = st ? 1 : 0;
        int const k = ld(tab.math, l);
        if (k < 0) continue;
        double re[3] = {0.0, 0.0, 0.0}, im[3] = {0.0, 0.0, 0.0};
        if (st) {
            double const dir = ld(tab.gen, l) ? 1.0 : -1.0;
            if (!ld(tab.asym, l)) {
                double const r0 = dir * p[0] / 1e6, i0 = dir * q[0] / 1e6;
                bool const bad = isnan(r0) || isnan(i0);
#pragma unroll
                for (int ph = 0; ph < 3; ++ph) {
                    re[ph] = bad ? __builtin_nan("") : r0;
                    im[ph] = bad ? __builtin_nan("") : i0;
                }
            } else {
#pragma unroll
                for (int ph = 0; ph < 3; ++ph) {
                    re[ph] = dir * p[ph] / (1e6 / 3.0);
                    im[ph] = dir * q[ph] / (1e6 / 3.0);
                }
                if (tab.symmetric) {
                    re[0] = (re[0] + re[1] + re[2]) / 3.0;
                    im[0] = (im[0] + im[1] + im[2]) / 3.0;
                }
            }
        }
        for (int ph = 0; ph < tab.P; ++ph) {
            s_inj[((size_t(k) * tab.P + ph) * 2) * sp + lane] = re[ph];
            s_inj[((size_t(k) * tab.P + ph) * 2 + 1) * sp + lane] = im[ph];
        }
    }
}

__global__ void init_state_kernel(PairState ps) {
    size_t const i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(ps.n)) return;
    ps.done[i] = ACTIVE;
    ps.err[i] = ERR_NONE;
    ps.iters[i] = 0;
    ps.devmax[i] = 0ull;
    ps.lastdev[i] = 0.0;
    if (ps.pert) {
        ps.pert[i] = 0;
        ps.rmax[i] = 0ull;
        ps.berr[i] = 0ull;
    }
}

__global__ void retire_linear_kernel(PairState ps, uint8_t const* __restrict__ pair_linear) {
    size_t const i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(ps.n) || !pair_linear[i] || ps.done[i] != ACTIVE) return;
    ps.done[i] = CONVERGED;
    ps.iters[i] = 1;
}

// pairs of padding lanes are never touched by the sweeps: their deviation stays 0 and they retire
// at the first check
__global__ void converge_kernel(PairState ps, int it_host, int const* __restrict__ it_dev, int max_iter, double tol,
                                int* active) {
    size_t const i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(ps.n)) return;
    if (ps.done[i] != ACTIVE) return;
    // a step from perturbed pivots waits for its refinement; the pair takes no part in the sweeps meanwhile
    if (ps.pert && ps.pert[i]) {
        ps.done[i] = REFINE;
        atomicAdd(active + 1, 1);
        return;
    }
    // every active pair iterates once per check: its own count (a pair that waited for a refinement is one
    // behind the launches)
    (void)it_host;
    (void)it_dev;
    int const it = ps.iters[i] + 1;
    double const md = __longlong_as_double(static_cast<long long>(ps.devmax[i]));
    ps.devmax[i] = 0ull;
    ps.iters[i] = it;
    if (!(md > tol)) {
        ps.done[i] = CONVERGED;
    } else if (it >= max_iter) {
        ps.done[i] = FAILED;
        ps.err[i] = ERR_DIVERGE;
        ps.lastdev[i] = md;
    } else {
        atomicAdd(active, 1);
    }
}

// start of a captured Newton-Raphson iteration: the iteration counter advances and the active count of
// this iteration's check starts at zero
__global__ void begin_iteration_kernel(int* it, int* active) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *it += 1;
        active[0] = 0;
        active[1] = 0;
    }
}

// ---------------------------------------------------------------------------------------------------
// iterative refinement of Newton-Raphson steps whose FP32 factorisation perturbed pivots
// (sparse_lu_solver.hpp:514-649: residual with the original matrix, componentwise backward error with the
// denominator capped below at 1e-4 of its largest value, correction solved with the perturbed factors)
// ---------------------------------------------------------------------------------------------------
__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* p, double v) {
    atomicMax(p, static_cast<unsigned long long>(__double_as_longlong(v)));
}
__device__ __forceinline__ double from_bits(unsigned long long b) { return __longlong_as_double(static_cast<long long>(b)); }

// the (group, bus, lane) of a flat index over every group's buses and lanes, when the lane's pair is in REFINE
struct RefineLane {
    int gr, bus, s, mm;
};
__device__ __forceinline__ bool refine_lane(DevGroup const* groups, long long const* offsets, int n_groups, RefineLane& r,
                                            DevGroup& g) {
    long long const gidx = (long long)(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gidx >= ld(offsets, n_groups)) return false;
    r.gr = find_group(offsets, n_groups, gidx);
    g = ld(groups, r.gr);
    size_t const sp = size_t(g.S_pad);
    size_t const idx = size_t(gidx - ld(offsets, r.gr));
    r.bus = int(idx / sp);
    r.s = int(idx % sp);
    if (r.s >= g.S || !g.xacc) return false;
    r.mm = ld(g.bus_mm, r.bus);
    return r.mm >= 0 && g.done[size_t(r.mm) * sp + r.s] == REFINE;
}

__global__ void refine_reset_kernel(PairState ps) {
    size_t const i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(ps.n) || ps.done[i] != REFINE) return;
    ps.rmax[i] = 0ull;
    ps.berr[i] = 0ull;
}

template <int P>
__global__ void refine_start_kernel(DevGroup const* __restrict__ groups, long long const* __restrict__ offsets,
                                    int n_groups) {
    RefineLane r;
    DevGroup g;
    if (!refine_lane(groups, offsets, n_groups, r, g)) return;
    size_t const sp = g.S_pad;
    // the step of the Newton-Raphson sweep, solved with the perturbed factors, starts the accumulation
#pragma unroll
    for (int c = 0; c < 2 * P; ++c) g.xacc[(size_t(r.bus) * 2 * P + c) * sp + r.s] = g.y[(size_t(r.bus) * 2 * P + c) * sp + r.s];
}

// residual r = rhs - J x of row `bus` at the present voltages (J applied in FP64, as assembled by the sweep
// before its FP32 rounding), into the forward vector; |rhs| + |J| |x| into rden and its largest value per pair
template <int P>
__global__ void refine_residual_kernel(DevGroup const* __restrict__ groups, long long const* __restrict__ offsets,
                                       int n_groups) {
    constexpr int NB = 2 * P;
    constexpr int NB2 = NB * NB;
    constexpr int PP = P * P;
    RefineLane r;
    DevGroup g;
    if (!refine_lane(groups, offsets, n_groups, r, g)) return;
    size_t const sp = g.S_pad;
    int const p = r.bus, s = r.s;
    int const rb = ld(g.row_ptr, p), re = ld(g.row_end, p);
    int const es = g.row_stride ? ld(g.row_stride, p) : 1;
    auto X = [&](int bus, int c) { return g.xacc[(size_t(bus) * NB + c) * sp + s]; };
    cplx up[P];
    load_u<P>(g, p, s, up);
    double D[NB2], rhs[NB], jx[NB], ajx[NB], pcal[P], qcal[P];
#pragma unroll
    for (int q = 0; q < NB2; ++q) D[q] = 0.0;
#pragma unroll
    for (int c = 0; c < NB; ++c) jx[c] = ajx[c] = 0.0;
#pragma unroll
    for (int ph = 0; ph < P; ++ph) pcal[ph] = qcal[ph] = 0.0;
    for (int k = rb; k < re; k += es) {
        int const j = ld(g.col, k);
        if (ld(g.fill, k)) continue;
        double yk[PP * 2], pr[P], qr[P];
        load_y<P, false>(g, k, s, yk);
        if (j == p) {
            hnml<P, double>(yk, up, up, D, pr, qr);
        } else {
            cplx uj[P];
            load_u<P>(g, j, s, uj);
            double a[NB2];
            hnml<P, double>(yk, up, uj, a, pr, qr);
#pragma unroll
            for (int rr = 0; rr < NB; ++rr)
#pragma unroll
                for (int c = 0; c < NB; ++c) {
                    double const xj = X(j, c);
                    jx[rr] += a[c * NB + rr] * xj;
                    ajx[rr] += fabs(a[c * NB + rr]) * fabs(xj);
                }
        }
#pragma unroll
        for (int ph = 0; ph < P; ++ph) {
            pcal[ph] += pr[ph];
            qcal[ph] += qr[ph];
        }
    }
    nr_bus_terms<P>(g, p, s, up, pcal, qcal, D, rhs);
    double den_max = 0.0;
#pragma unroll
    for (int rr = 0; rr < NB; ++rr) {
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            double const xp = X(p, c);
            jx[rr] += D[c * NB + rr] * xp;
            ajx[rr] += fabs(D[c * NB + rr]) * fabs(xp);
        }
        double const den = fabs(rhs[rr]) + ajx[rr];
        g.y[(size_t(p) * NB + rr) * sp + s] = rhs[rr] - jx[rr];
        g.rden[(size_t(p) * NB + rr) * sp + s] = den;
        den_max = den > den_max ? den : den_max;
    }
    atomic_max_nonneg(&g.rmax[size_t(r.mm) * sp + s], den_max);
}

// componentwise backward error of the residual, the denominators capped below at 1e-4 of the pair's largest
// (iterate_and_backward_error, sparse_lu_solver.hpp:582-626); per pair and over all pairs
template <int P>
__global__ void refine_berr_kernel(DevGroup const* __restrict__ groups, long long const* __restrict__ offsets,
                                   int n_groups, unsigned long long* __restrict__ gmax) {
    constexpr int NB = 2 * P;
    RefineLane r;
    DevGroup g;
    if (!refine_lane(groups, offsets, n_groups, r, g)) return;
    size_t const sp = g.S_pad;
    double const cap = 1e-4 * from_bits(g.rmax[size_t(r.mm) * sp + r.s]);
    double be = 0.0;
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        size_t const i = (size_t(r.bus) * NB + c) * sp + r.s;
        double const den = g.rden[i] > cap ? g.rden[i] : cap;
        double const e = den > 0.0 ? fabs(g.y[i]) / den : (g.y[i] == 0.0 ? 0.0 : __builtin_inf());
        be = e > be ? e : be;
    }
    atomic_max_nonneg(&g.berr[size_t(r.mm) * sp + r.s], be);
    atomic_max_nonneg(gmax, be);
}

// the refined step applied as the Newton-Raphson sweep applies a step (newton_raphson_pf_solver.hpp:325-349)
template <int P>
__global__ void refine_apply_kernel(DevGroup const* __restrict__ groups, long long const* __restrict__ offsets,
                                    int n_groups) {
    RefineLane r;
    DevGroup g;
    if (!refine_lane(groups, offsets, n_groups, r, g)) return;
    size_t const sp = g.S_pad;
    int const p = r.bus, s = r.s;
    double dbus = 0.0;
#pragma unroll
    for (int ph = 0; ph < P; ++ph) {
        size_t const b = (size_t(p) * P + ph) * 2;
        double const dth = g.xacc[(size_t(p) * 2 * P + ph) * sp + s];
        double const dv = g.xacc[(size_t(p) * 2 * P + ph + P) * sp + s];
        double const th0 = g.xv[b * sp + s], vm0 = g.xv[(b + 1) * sp + s];
        double const th = th0 + dth;
        double vm = vm0;
        vm += vm * dv;
        double sn, cs, sn0, cs0;
        sincos(th, &sn, &cs);
        sincos(th0, &sn0, &cs0);
        double const nr = vm * cs, ni = vm * sn;
        double const dr = nr - vm0 * cs0, di = ni - vm0 * sn0;
        double const d = sqrt(dr * dr + di * di);
        dbus = ph == 0 ? d : (d > dbus ? d : dbus);
        g.xv[b * sp + s] = th;
        g.xv[(b + 1) * sp + s] = vm;
        g.u[b * sp + s] = nr;
        g.u[(b + 1) * sp + s] = ni;
    }
    atomic_max_nonneg(&g.devmax[size_t(r.mm) * sp + s], dbus);
}

// per REFINE pair: a backward error still above the tolerance after the last pass fails the pair as the
// reference's refinement does (SparseMatrixError); otherwise the convergence decision of converge_kernel
__global__ void refine_finish_kernel(PairState ps, int max_iter, double tol, int* active) {
    size_t const i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(ps.n) || ps.done[i] != REFINE) return;
    bool const perturbed = ps.pert[i] >= 2;
    ps.pert[i] = 0;
    if (perturbed && !(from_bits(ps.berr[i]) <= kRefineTol)) {
        ps.done[i] = FAILED;
        ps.err[i] = ERR_SPARSE;
        return;
    }
    int const it = ps.iters[i] + 1;
    double const md = from_bits(ps.devmax[i]);
    ps.devmax[i] = 0ull;
    ps.iters[i] = it;
    if (!(md > tol)) {
        ps.done[i] = CONVERGED;
    } else if (it >= max_iter) {
        ps.done[i] = FAILED;
        ps.err[i] = ERR_DIVERGE;
        ps.lastdev[i] = md;
    } else {
        ps.done[i] = ACTIVE;
        atomicAdd(active, 1);
    }
}
__global__ void set_int_kernel(int* p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

template <int P>
__global__ void ic_start_kernel(DevGroup const* __restrict__ groups, long long const* __restrict__ offsets,
                                int n_groups) {
    long long const gidx = (long long)(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gidx >= offsets[n_groups]) return;
    int const gr = find_group(offsets, n_groups, gidx);
    DevGroup const g = ld(groups, gr);
    size_t const sp = g.S_pad;
    size_t const idx = size_t(gidx - offsets[gr]);
    int const b = int(idx / sp), s = int(idx % sp);
    if (s >= g.S) return;
    int const mm = ld(g.bus_mm, b);
    if (mm < 0 || g.done[size_t(mm) * sp + s] != ACTIVE) return;
    int const q0 = ld(g.mm_src_ptr, mm), q1 = ld(g.mm_src_ptr, mm + 1);
    cplx sum = {0.0, 0.0};
    for (int qi = q0; qi < q1; ++qi) {
        int const q = ld(g.mm_src, qi);
        cplx const ur = {g.u_src[(size_t(q) * 2) * sp + s], g.u_src[(size_t(q) * 2 + 1) * sp + s]};
        double sn, cs;
        sincos(-ld(g.phase, g.src_bus[q]), &sn, &cs);
        sum = cadd(sum, cmul(ur, {cs, sn}));
    }
    double const inv = 1.0 / double(q1 - q0);
    double sn, cs;
    sincos(ld(g.phase, b), &sn, &cs);
    cplx ub[P];
    balanced<P>(cmul({sum.re * inv, sum.im * inv}, {cs, sn}), ub);
#pragma unroll
    for (int ph = 0; ph < P; ++ph) {
        g.u[((size_t(b) * P + ph) * 2) * sp + s] = ub[ph].re;
        g.u[((size_t(b) * P + ph) * 2 + 1) * sp + s] = ub[ph].im;
    }
}

// ---------------------------------------------------------------------------------------------------
// results per bus: load powers and currents, bus injection and source results, shunt flows
// ---------------------------------------------------------------------------------------------------
template <int P>
__device__ __forceinline__ void bus_result(DevOutput const& o, int b, int s) {
    size_t const sp = o.S_pad;
    if (s >= o.S || !o.ok[s] || ld(o.bus_mm, b) < 0) return;
    if (o.dead && o.dead[size_t(b) * sp + s]) {
        // not energised in this lane: the appliances of the bus carry nothing
        for (int q = ld(o.shunt_ptr, b); q < ld(o.shunt_ptr, b + 1); ++q)
            for (int c = 0; c < 4 * P; ++c) o.shunt_si[(size_t(q) * 4 * P + c) * sp + s] = 0.0;
        for (int c = 0; c < 2 * P; ++c) o.bus_inj[(size_t(b) * 2 * P + c) * sp + s] = 0.0;
        return;
    }
    auto U = [&](int bus, int ph) {
        return cplx{o.u[((size_t(bus) * P + ph) * 2) * sp + s], o.u[((size_t(bus) * P + ph) * 2 + 1) * sp + s]};
    };
    cplx ub[P];
#pragma unroll
    for (int ph = 0; ph < P; ++ph) ub[ph] = U(b, ph);
    cplx iload[P], acc_s[P];
#pragma unroll
    for (int ph = 0; ph < P; ++ph) iload[ph] = acc_s[ph] = {0.0, 0.0};
    for (int l = ld(o.load_ptr, b); l < ld(o.load_ptr, b + 1); ++l) {
        int const type = o.linear ? 1 : ld(o.load_type, l);
#pragma unroll
        for (int ph = 0; ph < P; ++ph) {
            cplx sv = {o.s_inj[((size_t(l) * P + ph) * 2) * sp + s], o.s_inj[((size_t(l) * P + ph) * 2 + 1) * sp + s]};
            double const a2 = ub[ph].re * ub[ph].re + ub[ph].im * ub[ph].im;
            double const f = type == 0 ? 1.0 : (type == 1 ? a2 : sqrt(a2));
            sv = {sv.re * f, sv.im * f};
            cplx const iv = cconj(cdiv(sv, ub[ph]));
            iload[ph] = cadd(iload[ph], iv);
            acc_s[ph] = cadd(acc_s[ph], sv);
        }
    }
    int const s0 = ld(o.src_ptr, b), s1 = ld(o.src_ptr, b + 1);
    if (s1 > s0) {
        // bus injection u * conj(sum_j Y_bj u_j)
        cplx inj[P];
#pragma unroll
        for (int ph = 0; ph < P; ++ph) inj[ph] = {0.0, 0.0};
        int const es = o.row_stride ? ld(o.row_stride, b) : 1;
        for (int k = ld(o.row_ptr, b); k < ld(o.row_end, b); k += es) {
            if (ld(o.fill, k)) continue;
            int const j = ld(o.col, k);
            double y[P * P * 2];
#pragma unroll
            for (int q = 0; q < P * P * 2; ++q)
                y[q] = o.ylane ? o.ylane[(size_t(k) * P * P * 2 + q) * sp + s] : ld(o.ydata, size_t(k) * P * P * 2 + q);
#pragma unroll
            for (int r = 0; r < P; ++r)
#pragma unroll
                for (int c = 0; c < P; ++c)
                    inj[r] = cadd(inj[r], cmul({y[(r * P + c) * 2], y[(r * P + c) * 2 + 1]}, U(j, c)));
        }
        cplx it_[P];
#pragma unroll
        for (int ph = 0; ph < P; ++ph) {
            cplx const sb = cmul(cconj(inj[ph]), ub[ph]);
            it_[ph] = csub(cconj(cdiv(sb, ub[ph])), iload[ph]);
        }
        auto put_src = [&](int q, cplx const (&iv)[P]) {
#pragma unroll
            for (int ph = 0; ph < P; ++ph) {
                cplx const sv = cmul(ub[ph], cconj(iv[ph]));
                size_t const base = (size_t(q) * 2) * P;
                o.src_si[((base + ph) * 2) * sp + s] = sv.re;
                o.src_si[((base + ph) * 2 + 1) * sp + s] = sv.im;
                o.src_si[((base + P + ph) * 2) * sp + s] = iv[ph].re;
                o.src_si[((base + P + ph) * 2 + 1) * sp + s] = iv[ph].im;
                acc_s[ph] = cadd(acc_s[ph], sv);
            }
        };
        auto ysrc = [&](int q, int k) {
            return cplx{o.y_src[(size_t(q) * 4 + 2 * k) * sp + s], o.y_src[(size_t(q) * 4 + 2 * k + 1) * sp + s]};
        };
        auto usrc = [&](int q) { return cplx{o.u_src[(size_t(q) * 2) * sp + s], o.u_src[(size_t(q) * 2 + 1) * sp + s]}; };
        if (s1 - s0 == 1) {
            put_src(s0, it_);
        } else if (P == 1) {
            cplx yt = {0, 0}, irt = {0, 0};
            for (int q = s0; q < s1; ++q) yt = cadd(yt, ysrc(q, 0));
            cplx const zt = cdiv({1.0, 0.0}, yt);
            for (int q = s0; q < s1; ++q) irt = cadd(irt, cmul(usrc(q), ysrc(q, 0)));
            for (int q = s0; q < s1; ++q) {
                cplx const ratio = cmul(ysrc(q, 0), zt);
                cplx const lhs = cmul(ratio, csub(cmul(usrc(q), yt), irt));
                cplx iv[P];
                iv[0] = cadd(lhs, cmul(ratio, it_[0]));
                put_src(q, iv);
            }
        } else {
            cplx const a = {-0.5, 0.86602540378443864676}, a2 = {-0.5, -0.86602540378443864676};
            cplx yt0 = {0, 0}, yt1 = {0, 0}, ir1 = {0, 0};
            for (int q = s0; q < s1; ++q) {
                yt0 = cadd(yt0, ysrc(q, 1));
                yt1 = cadd(yt1, ysrc(q, 0));
                ir1 = cadd(ir1, cmul(usrc(q), ysrc(q, 0)));
            }
            // to sequence components: (1/3) [[1,1,1],[1,a,a2],[1,a2,a]]
            cplx const i0 = {(it_[0].re + it_[1 % P].re + it_[2 % P].re) / 3.0,
                             (it_[0].im + it_[1 % P].im + it_[2 % P].im) / 3.0};
            cplx const i1t = cadd(cadd(it_[0], cmul(a, it_[1 % P])), cmul(a2, it_[2 % P]));
            cplx const i2t = cadd(cadd(it_[0], cmul(a2, it_[1 % P])), cmul(a, it_[2 % P]));
            cplx const i1 = {i1t.re / 3.0, i1t.im / 3.0}, i2 = {i2t.re / 3.0, i2t.im / 3.0};
            for (int q = s0; q < s1; ++q) {
                cplx const r1 = cdiv(ysrc(q, 0), yt1);
                cplx const lhs = cmul(r1, csub(cmul(usrc(q), yt1), ir1));
                cplx const v0 = cmul(cdiv(ysrc(q, 1), yt0), i0);
                cplx const v1 = cadd(lhs, cmul(r1, i1));
                cplx const v2 = cmul(cdiv(ysrc(q, 0), yt1), i2);
                // back to phases: [[1,1,1],[1,a2,a],[1,a,a2]]
                cplx iv[P];
                iv[0] = cadd(cadd(v0, v1), v2);
                iv[1 % P] = cadd(cadd(v0, cmul(a2, v1)), cmul(a, v2));
                iv[2 % P] = cadd(cadd(v0, cmul(a, v1)), cmul(a2, v2));
                put_src(q, iv);
            }
        }
    }
#pragma unroll
    for (int ph = 0; ph < P; ++ph) {
        o.bus_inj[((size_t(b) * P + ph) * 2) * sp + s] = acc_s[ph].re;
        o.bus_inj[((size_t(b) * P + ph) * 2 + 1) * sp + s] = acc_s[ph].im;
    }
    for (int q = ld(o.shunt_ptr, b); q < ld(o.shunt_ptr, b + 1); ++q) {
        double const* y = o.shunt_y + size_t(q) * P * P * 2;
#pragma unroll
        for (int r = 0; r < P; ++r) {
            cplx iy = {0.0, 0.0};
#pragma unroll
            for (int c = 0; c < P; ++c) iy = cadd(iy, cmul({y[(r * P + c) * 2], y[(r * P + c) * 2 + 1]}, ub[c]));
            cplx const iv = {-iy.re, -iy.im};
            cplx const sv = cmul(ub[r], cconj(iv));
            size_t const base = (size_t(q) * 2) * P;
            o.shunt_si[((base + r) * 2) * sp + s] = sv.re;
            o.shunt_si[((base + r) * 2 + 1) * sp + s] = sv.im;
            o.shunt_si[((base + P + r) * 2) * sp + s] = iv.re;
            o.shunt_si[((base + P + r) * 2 + 1) * sp + s] = iv.im;
        }
    }
}

// one thread per (bus, lane); a wave whose 64 threads are 64 lanes of one bus (groups of 64 lanes
// and more) takes the bus from wave-uniform values, so the bus's structure is read with scalar loads
template <int P>
__global__ void bus_results_kernel(DevOutput const* __restrict__ outs, long long const* __restrict__ offsets,
                                   int n_groups) {
    long long const total = ld(offsets, n_groups);
    long long const wbase = (long long)(blockIdx.x) * blockDim.x +
                            (long long)(__builtin_amdgcn_readfirstlane(int(threadIdx.x) / kWave)) * kWave;
    if (wbase >= total) return;
    int const lane = int(threadIdx.x) % kWave;
    {
        int const gr = find_group(offsets, n_groups, wbase);
        DevOutput const o = ld(outs, gr);
        size_t const sp = o.S_pad;
        size_t const idx0 = size_t(wbase - ld(offsets, gr));
        size_t const s0 = idx0 % sp;
        if (sp % kWave == 0 && s0 + kWave <= sp) {
            bus_result<P>(o, int(idx0 / sp), int(s0) + lane);
            return;
        }
    }
    long long const gidx = wbase + lane;
    if (gidx >= total) return;
    int const gr = find_group(offsets, n_groups, gidx);
    DevOutput const o = ld(outs, gr);
    size_t const idx = size_t(gidx - ld(offsets, gr));
    bus_result<P>(o, int(idx / o.S_pad), int(idx % o.S_pad));
}

// ---------------------------------------------------------------------------------------------------
// record writers (node.hpp, branch.hpp, appliance.hpp get_output)
// ---------------------------------------------------------------------------------------------------
// one record for component i of lane s (the lane is valid and converged)
template <int P>
__device__ __forceinline__ void node_record(DevOutput const& o, NodeRec const& r, int i, int s, NodeOutput<P>& rec,
                                            bool bus_sums) {
    size_t const sp = o.S_pad;
    rec.id = ld(r.id, i);
    int const bus = ld(r.bus, i);
    double const base_power = P == 1 ? 1e6 : 1e6 / 3.0;
    double const u_scale = P == 1 ? 1.0 : 0.57735026918962576451;
    if (bus >= 0 && !(o.dead && ld(o.dead, size_t(bus) * sp + s))) {
        rec.energized = 1;
#pragma unroll
        for (int ph = 0; ph < P; ++ph) {
            double const ur = o.u[((size_t(bus) * P + ph) * 2) * sp + s];
            double const ui = o.u[((size_t(bus) * P + ph) * 2 + 1) * sp + s];
            rec.u_pu[ph] = sqrt(ur * ur + ui * ui);
            rec.u[ph] = u_scale * ld(r.u_rated, i) * rec.u_pu[ph];
            rec.u_angle[ph] = atan2(ui, ur);
            if (bus_sums) {
                rec.p[ph] = base_power * o.bus_inj[((size_t(bus) * P + ph) * 2) * sp + s];
                rec.q[ph] = base_power * o.bus_inj[((size_t(bus) * P + ph) * 2 + 1) * sp + s];
            }
        }
    }
}

template <int P>
__device__ __forceinline__ void branch_record(DevOutput const& o, BranchRec const& r, int i, int s, BranchOutput<P>& rec) {
    size_t const sp = o.S_pad;
    rec.id = ld(r.id, i);
    int const m = ld(r.math, i);
    double const base_power = P == 1 ? 1e6 : 1e6 / 3.0;
    int f = m >= 0 ? ld(r.from_bus, m) : -1, t = m >= 0 ? ld(r.to_bus, m) : -1;
    double const* y = m >= 0 ? r.branch_y + size_t(m) * 4 * P * P * 2 : nullptr;
    bool energized = m >= 0;
    if (m >= 0 && r.lane_status) {
        // union group: the side statuses of this lane pick the admittance; energised when a closed
        // side meets a live bus
        int const stt = ld(r.lane_status, size_t(m) * sp + s);
        if (!(stt & 1)) f = -1;
        if (!(stt & 2)) t = -1;
        energized = (f >= 0 && !ld(o.dead, size_t(f) * sp + s)) || (t >= 0 && !ld(o.dead, size_t(t) * sp + s));
        y = r.branch_y_var + (size_t(m) * 4 + stt) * 4 * P * P * 2;
    }
    if (energized) {
        rec.energized = 1;
        cplx uf[P], ut[P];
#pragma unroll
        for (int ph = 0; ph < P; ++ph) {
            uf[ph] = f >= 0 ? cplx{o.u[((size_t(f) * P + ph) * 2) * sp + s], o.u[((size_t(f) * P + ph) * 2 + 1) * sp + s]}
                            : cplx{0.0, 0.0};
            ut[ph] = t >= 0 ? cplx{o.u[((size_t(t) * P + ph) * 2) * sp + s], o.u[((size_t(t) * P + ph) * 2 + 1) * sp + s]}
                            : cplx{0.0, 0.0};
        }
        auto Y = [&](int w, int rr, int cc) {
            return cplx{ld(y, ((w * P + rr) * P + cc) * 2), ld(y, ((w * P + rr) * P + cc) * 2 + 1)};
        };
        double sum_sf = 0, sum_st = 0, max_i = 0;
#pragma unroll
        for (int ph = 0; ph < P; ++ph) {
            cplx ifr = {0, 0}, ito = {0, 0};
#pragma unroll
            for (int c = 0; c < P; ++c) {
                ifr = cadd(ifr, cadd(cmul(Y(0, ph, c), uf[c]), cmul(Y(1, ph, c), ut[c])));
                ito = cadd(ito, cadd(cmul(Y(2, ph, c), uf[c]), cmul(Y(3, ph, c), ut[c])));
            }
            cplx const sf = cmul(uf[ph], cconj(ifr));
            cplx const st = cmul(ut[ph], cconj(ito));
            rec.p_from[ph] = base_power * sf.re;
            rec.q_from[ph] = base_power * sf.im;
            rec.i_from[ph] = ld(r.base_i_from, i) * cabs_(ifr);
            rec.s_from[ph] = base_power * cabs_(sf);
            rec.p_to[ph] = base_power * st.re;
            rec.q_to[ph] = base_power * st.im;
            rec.i_to[ph] = ld(r.base_i_to, i) * cabs_(ito);
            rec.s_to[ph] = base_power * cabs_(st);
            sum_sf += rec.s_from[ph];
            sum_st += rec.s_to[ph];
            double const mi = rec.i_from[ph] > rec.i_to[ph] ? rec.i_from[ph] : rec.i_to[ph];
            max_i = ph == 0 ? mi : (mi > max_i ? mi : max_i);
        }
        rec.loading = r.is_line ? max_i / ld(r.rating, i) : (sum_sf > sum_st ? sum_sf : sum_st) / ld(r.rating, i);
    }
}

template <int P>
__device__ __forceinline__ void appliance_record(DevOutput const& o, ApplRec const& r, int i, int s, ApplianceOutput<P>& rec) {
    size_t const sp = o.S_pad;
    rec.id = ld(r.id, i);
    int const m = ld(r.math, i);
    double const base_power = P == 1 ? 1e6 : 1e6 / 3.0;
    int const bus = ld(r.bus, i);
    bool const dead = o.dead && bus >= 0 && ld(o.dead, size_t(bus) * sp + s);
    if (m >= 0 && !dead && (r.si || bus >= 0)) {
        rec.energized = r.status ? ld(r.status, size_t(i) * sp + s) : ld(r.status_const, i);
        size_t const base = (size_t(m) * 2) * P;
        // load_gens: s = s_inj f(|u|) by type (1, |u|^2, |u| for constant power, impedance, current; the
        // linear method treats every load as constant impedance), i = conj(s / u)
        int const type = r.si ? 0 : (o.linear ? 1 : ld(o.load_type, m));
#pragma unroll
        for (int ph = 0; ph < P; ++ph) {
            cplx sv, iv;
            if (r.si) {
                sv = {r.si[((base + ph) * 2) * sp + s], r.si[((base + ph) * 2 + 1) * sp + s]};
                iv = {r.si[((base + P + ph) * 2) * sp + s], r.si[((base + P + ph) * 2 + 1) * sp + s]};
            } else {
                cplx const ub = {o.u[((size_t(bus) * P + ph) * 2) * sp + s], o.u[((size_t(bus) * P + ph) * 2 + 1) * sp + s]};
                sv = {o.s_inj[((size_t(m) * P + ph) * 2) * sp + s], o.s_inj[((size_t(m) * P + ph) * 2 + 1) * sp + s]};
                double const a2 = ub.re * ub.re + ub.im * ub.im;
                double const f = type == 0 ? 1.0 : (type == 1 ? a2 : sqrt(a2));
                sv = {sv.re * f, sv.im * f};
                iv = cconj(cdiv(sv, ub));
            }
            rec.p[ph] = base_power * sv.re * r.direction;
            rec.q[ph] = base_power * sv.im * r.direction;
            rec.s[ph] = base_power * cabs_(sv);
            rec.i[ph] = ld(r.base_i, i) * cabs_(iv);
            rec.pf[ph] = rec.s[ph] < 1e-8 ? 0.0 : rec.p[ph] / rec.s[ph];
        }
    }
}

template <int P>
__device__ __forceinline__ void make_record(DevOutput const& o, NodeRec const& r, int i, int s, NodeOutput<P>& rec,
                                            bool bus_sums) {
    node_record<P>(o, r, i, s, rec, bus_sums);
}
template <int P>
__device__ __forceinline__ void make_record(DevOutput const& o, BranchRec const& r, int i, int s, BranchOutput<P>& rec,
                                            bool) {
    branch_record<P>(o, r, i, s, rec);
}
template <int P>
__device__ __forceinline__ void make_record(DevOutput const& o, ApplRec const& r, int i, int s,
                                            ApplianceOutput<P>& rec, bool) {
    appliance_record<P>(o, r, i, s, rec);
}

// A block builds a tile of W lanes x (64 TN / W) components in LDS (W = min(S_pad, 64)), reading the
// scenario-interleaved state with lane-contiguous (coalesced) loads, then writes each lane's records as
// one contiguous run of 16-byte stores: both sides of the transpose stay coalesced, for wide groups and
// for single-scenario ones. Tile t of group g: offsets[g] <= t.
// With columns (cols.n > 0) only the listed attributes leave the tile, each into its own column: the
// lanes of a tile row write consecutive elements of one scenario's column.
// Workgroups are handed to the chip's 8 XCDs in turn. Tile b of the grid takes tile xcd_tile(b) instead,
// which gives every XCD one contiguous run of tiles: neighbouring tiles write neighbouring pieces of the
// same rows (columns and records), and those partial lines then meet in one L2. A bijection of [0, nb).
__device__ __forceinline__ long long xcd_tile(long long b, long long nb) {
    constexpr long long kXcd = 8;
    long long const x = b % kXcd, q = b / kXcd, base = nb / kXcd, extra = nb % kXcd;
    return x * base + (x < extra ? x : extra) + q;
}

template <int P, class Rec, class Out>
__global__ __launch_bounds__(256) void records_kernel(DevOutput const* __restrict__ outs, Rec const* __restrict__ recs,
                                                      long long const* __restrict__ tile_off, int n_groups,
                                                      Out* __restrict__ out, ColumnSet const cols) {
    constexpr int SLOTS = kWaveLanes * record_tile(sizeof(Out));
    constexpr int CH = int(sizeof(Out) / 16);
    static_assert(sizeof(Out) % 16 == 0, "records are copied in 16-byte chunks");
    __shared__ __align__(16) Out tile[SLOTS];
    long long const t = xcd_tile(blockIdx.x, gridDim.x);
    int const gr = find_group(tile_off, n_groups, t);
    DevOutput const o = ld(outs, gr);
    Rec const r = ld(recs, gr);
    int const W = o.S_pad < kWaveLanes ? o.S_pad : kWaveLanes;
    int const TNg = SLOTS / W;
    int const n_itiles = (r.n_comp + TNg - 1) / TNg;
    long long const local = t - tile_off[gr];
    int const s0 = int(local / n_itiles) * W, i0 = int(local % n_itiles) * TNg;
    auto fill = [&](int lane, int ii) {
        int const s = s0 + lane, i = i0 + ii;
        if (s < o.S && o.ok[s] && i < r.n_comp) {
            // zero the slot (padding included: the records come out byte for byte the same every run),
            // then fill the attributes in place
            Out& slot = tile[lane * TNg + ii];
            uint4* z = reinterpret_cast<uint4*>(&slot);
#pragma unroll
            for (int k = 0; k < CH; ++k) z[k] = uint4{0u, 0u, 0u, 0u};
            make_record<P>(o, r, i, s, slot, cols.bus_sums != 0);
        }
    };
    if (W == kWaveLanes) {
        // a wave fills one component for 64 lanes: the component's data are wave-uniform (scalar loads)
        for (int q = int(threadIdx.x); q < W * TNg; q += int(blockDim.x))
            fill(q & (kWaveLanes - 1), __builtin_amdgcn_readfirstlane(q / kWaveLanes));
    } else {
        for (int q = int(threadIdx.x); q < W * TNg; q += int(blockDim.x)) fill(q % W, q / W);
    }
    __syncthreads();
    int const n_valid = r.n_comp - i0 < TNg ? r.n_comp - i0 : TNg;
    if (cols.n > 0) {
        for (int a = 0; a < cols.n; ++a) {
            int const off = cols.off[a], sz = cols.size[a];
            for (int c = int(threadIdx.x); c < W * TNg; c += int(blockDim.x)) {
                int const ln = c / TNg, k = c % TNg;
                int const sl = s0 + ln;
                if (k >= n_valid || sl >= o.S || !o.ok[sl]) continue;
                char const* src = reinterpret_cast<char const*>(&tile[ln * TNg + k]) + off;
                char* dst = cols.dst[a] + (size_t(ld(o.scenario_index, sl)) * r.n_comp + i0 + k) * size_t(sz);
                if (sz % 8 == 0) {
                    for (int q = 0; q < sz / 8; ++q)
                        reinterpret_cast<double*>(dst)[q] = reinterpret_cast<double const*>(src)[q];
                } else if (sz == 4) {
                    *reinterpret_cast<int*>(dst) = *reinterpret_cast<int const*>(src);
                } else {
                    for (int q = 0; q < sz; ++q) dst[q] = src[q];
                }
            }
        }
        return;
    }
    for (int c = int(threadIdx.x); c < W * TNg * CH; c += int(blockDim.x)) {
        int const ln = c / (TNg * CH), k = c % (TNg * CH);
        int const sl = s0 + ln;
        if (k >= n_valid * CH || sl >= o.S || !o.ok[sl]) continue;
        uint4 const* src = reinterpret_cast<uint4 const*>(&tile[ln * TNg]) + k;
        uint4* dst = reinterpret_cast<uint4*>(out + size_t(ld(o.scenario_index, sl)) * r.n_comp + i0) + k;
        *dst = *src;
    }
}

template <int P, class FT, int KIND>
void launch_fwd(DevGroup const* groups, WorkItem const* items, int n_items, int mode, hipStream_t st) {
    dim3 const grid{unsigned(n_items)};
    if (mode == MODE_INIT)
        hipLaunchKernelGGL((forward_kernel<P, FT, MODE_INIT, KIND>), grid, dim3(kWave), 0, st, groups, items);
    else if (mode == MODE_NR)
        hipLaunchKernelGGL((forward_kernel<P, FT, MODE_NR, KIND>), grid, dim3(kWave), 0, st, groups, items);
    else if (mode == MODE_ICF)
        hipLaunchKernelGGL((forward_kernel<P, FT, MODE_ICF, KIND>), grid, dim3(kWave), 0, st, groups, items);
    else if (mode == MODE_RESID)
        hipLaunchKernelGGL((forward_kernel<P, FT, MODE_RESID, KIND>), grid, dim3(kWave), 0, st, groups, items);
    else if (mode == MODE_REFINE) {
        if constexpr (sizeof(FT) == 4 && KIND != KIND_UNION)
            hipLaunchKernelGGL((forward_kernel<P, FT, MODE_REFINE, KIND>), grid, dim3(kWave), 0, st, groups, items);
    }
    else if constexpr (KIND == KIND_PLAIN) {
        if (mode == MODE_NRA)
            hipLaunchKernelGGL((forward_kernel<P, FT, MODE_NRA, KIND_PLAIN>), grid, dim3(kWave), 0, st, groups, items);
        else
            hipLaunchKernelGGL((forward_kernel<P, FT, MODE_RAW, KIND_PLAIN>), grid, dim3(kWave), 0, st, groups, items);
    }
}
template <int P, class FT, int KIND>
void launch_bwd(DevGroup const* groups, WorkItem const* items, int n_items, int mode, hipStream_t st) {
    dim3 const grid{unsigned(n_items)};
    if (mode == MODE_INIT)
        hipLaunchKernelGGL((backward_kernel<P, FT, MODE_INIT, KIND>), grid, dim3(kWave), 0, st, groups, items);
    else if (mode == MODE_NR)
        hipLaunchKernelGGL((backward_kernel<P, FT, MODE_NR, KIND>), grid, dim3(kWave), 0, st, groups, items);
    else if (mode == MODE_ICF || mode == MODE_RESID)
        hipLaunchKernelGGL((backward_kernel<P, FT, MODE_RESID, KIND>), grid, dim3(kWave), 0, st, groups, items);
    else if (mode == MODE_REFINE) {
        if constexpr (sizeof(FT) == 4 && KIND != KIND_UNION)
            hipLaunchKernelGGL((backward_kernel<P, FT, MODE_REFINE, KIND>), grid, dim3(kWave), 0, st, groups, items);
    }
    else if constexpr (KIND == KIND_PLAIN)
        hipLaunchKernelGGL((backward_kernel<P, FT, MODE_RAW, KIND_PLAIN>), grid, dim3(kWave), 0, st, groups, items);
}
template <bool FWD, int KIND>
void launch_sweep(DevGroup const* groups, WorkItem const* items, int n_items, int P, bool fp32, int mode,
                  hipStream_t st) {
    if (P == 1) {
        if constexpr (FWD)
            fp32 ? launch_fwd<1, float, KIND>(groups, items, n_items, mode, st)
                 : launch_fwd<1, double, KIND>(groups, items, n_items, mode, st);
        else
            fp32 ? launch_bwd<1, float, KIND>(groups, items, n_items, mode, st)
                 : launch_bwd<1, double, KIND>(groups, items, n_items, mode, st);
    } else {
        if constexpr (FWD)
            fp32 ? launch_fwd<3, float, KIND>(groups, items, n_items, mode, st)
                 : launch_fwd<3, double, KIND>(groups, items, n_items, mode, st);
        else
            fp32 ? launch_bwd<3, float, KIND>(groups, items, n_items, mode, st)
                 : launch_bwd<3, double, KIND>(groups, items, n_items, mode, st);
    }
}
template <bool FWD>
void launch_kind(DevGroup const* groups, WorkItem const* items, int n_items, int P, bool fp32, int mode,
                 hipStream_t st, int kind) {
    if (n_items <= 0) return;
    if (kind == KIND_UNION)
        launch_sweep<FWD, KIND_UNION>(groups, items, n_items, P, fp32, mode, st);
    else if (kind == KIND_PACKED)
        launch_sweep<FWD, KIND_PACKED>(groups, items, n_items, P, fp32, mode, st);
    else
        launch_sweep<FWD, KIND_PLAIN>(groups, items, n_items, P, fp32, mode, st);
}

__global__ void broadcast_y_kernel(double const* __restrict__ ydata, long long n_val, int S_pad,
                                   double* __restrict__ ylane) {
    long long const i = (long long)(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n_val * S_pad) return;
    ylane[i] = ydata[i / S_pad];
}
// ids of an update buffer in HBM: bit 0 when any id is NA, bit 1 when any is a number, bit 2 when some
// scenario's id differs from scenario 0's at the same position. eps > 0: rows row0 + s eps + e of n_scn
// scenarios, a thread per element e walking scenarios blockIdx.y, + gridDim.y, ... (scenario 0's id read
// once); eps <= 0: every row once, no comparison
__global__ void id_scan_kernel(char const* __restrict__ base, long long stride, long long total, long long row0,
                               long long eps, long long n_scn, int* __restrict__ flags) {
    auto id_at = [&](long long i) { return *reinterpret_cast<int const*>(base + i * stride); };
    int f = 0;
    if (eps > 0) {
        long long const e = (long long)(blockIdx.x) * blockDim.x + threadIdx.x;
        if (e < eps) {
            int const id0 = id_at(row0 + e);
            for (long long sc = blockIdx.y; sc < n_scn; sc += gridDim.y) {
                int const id = id_at(row0 + sc * eps + e);
                f |= id == na_IntID ? 1 : 2;
                if (id != id0) f |= 4;
            }
        }
    } else {
        long long const i = (long long)(blockIdx.x) * blockDim.x + threadIdx.x;
        if (i < total) f = id_at(i) == na_IntID ? 1 : 2;
    }
    // one atomic per wave
    for (int off = 32; off > 0; off >>= 1) f |= __shfl_xor(f, off);
    if ((threadIdx.x & 63) == 0 && f) atomicOr(flags, f);
}
__global__ void expand_member_rows_kernel(uint8_t const* __restrict__ src, int const* __restrict__ lane_member,
                                          int cols, int S_pad, uint8_t* __restrict__ dst) {
    long long const i = (long long)(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= (long long)(cols) * S_pad) return;
    int const c = int(i / S_pad), lane = int(i % S_pad);
    int const m = lane_member[lane];
    dst[i] = m < 0 ? uint8_t(0) : src[(long long)(m) * cols + c];
}
__global__ void patch_y_kernel(int n_patch, int Q, int S_pad, int const* __restrict__ lane,
                               int const* __restrict__ entry, double const* __restrict__ val,
                               double* __restrict__ ylane) {
    long long const i = (long long)(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= (long long)(n_patch) * Q) return;
    int const q = int(i / Q), c = int(i % Q);
    ylane[(size_t(entry[q]) * Q + c) * S_pad + lane[q]] = val[i];
}

__global__ void copy_probe_kernel(float const* __restrict__ a, float* __restrict__ b, long long n) {
    long long const i = (long long)(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

}  // namespace dev

using namespace dev;

int copy_probe(long long n) {
    if (n <= 0) return -1;
    float *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, size_t(n) * sizeof(float)) != hipSuccess) return -1;
    if (hipMalloc(&b, size_t(n) * sizeof(float)) != hipSuccess) {
        (void)hipFree(a);
        return -1;
    }
    int rc = hipMemset(a, 0, size_t(n) * sizeof(float)) == hipSuccess ? 0 : -1;
    if (rc == 0) {
        hipLaunchKernelGGL(copy_probe_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, nullptr, a, b, n);
        rc = (hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess) ? 0 : -1;
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

// 4x4 blocks of the state-estimation gain matrix (P = 2 gives NB = 4): factorise and solve raw systems,
// and the refinement sweeps with the stored factors
void launch_gain_sweep(DevGroup const* groups, WorkItem const* items, int n_items, bool refine, bool fwd,
                       hipStream_t st) {
    if (n_items == 0) return;
    dim3 const grid{unsigned(n_items)};
    if (fwd) {
        if (refine)
            hipLaunchKernelGGL((dev::forward_kernel<2, double, MODE_REFINE, KIND_PLAIN>), grid, dim3(dev::kWave), 0, st, groups,
                               items);
        else
            hipLaunchKernelGGL((dev::forward_kernel<2, double, MODE_RAWP, KIND_PLAIN>), grid, dim3(dev::kWave), 0, st, groups,
                               items);
    } else {
        if (refine)
            hipLaunchKernelGGL((dev::backward_kernel<2, double, MODE_REFINE, KIND_PLAIN>), grid, dim3(dev::kWave), 0, st, groups,
                               items);
        else
            hipLaunchKernelGGL((dev::backward_kernel<2, double, MODE_RAW, KIND_PLAIN>), grid, dim3(dev::kWave), 0, st, groups,
                               items);
    }
}

void launch_forward(DevGroup const* groups, WorkItem const* items, int n_items, int P, bool fp32, int mode,
                    hipStream_t st, int kind) {
    launch_kind<true>(groups, items, n_items, P, fp32, mode, st, kind);
}
void launch_backward(DevGroup const* groups, WorkItem const* items, int n_items, int P, bool fp32, int mode,
                     hipStream_t st, int kind) {
    launch_kind<false>(groups, items, n_items, P, fp32, mode, st, kind);
}
void launch_id_scan(char const* base, long long stride, long long total, long long row0, long long eps, int* flags,
                    hipStream_t st) {
    if (total <= 0) return;
    if (eps > 0) {
        long long const n_scn = (total - row0) / eps;
        dim3 const grid(unsigned((eps + 255) / 256), unsigned(std::min<long long>(std::max<long long>(n_scn, 1), 1024)));
        hipLaunchKernelGGL(id_scan_kernel, grid, dim3(256), 0, st, base, stride, total, row0, eps, n_scn, flags);
        // rows outside the scenarios' range (before row0 or after the last scenario) are scanned alone
        if (row0 > 0)
            hipLaunchKernelGGL(id_scan_kernel, dim3(unsigned((row0 + 255) / 256)), dim3(256), 0, st, base, stride, row0,
                               0LL, 0LL, 0LL, flags);
        long long const end = row0 + n_scn * eps;
        if (end < total)
            hipLaunchKernelGGL(id_scan_kernel, dim3(unsigned((total - end + 255) / 256)), dim3(256), 0, st,
                               base + end * stride, stride, total - end, 0LL, 0LL, 0LL, flags);
    } else {
        hipLaunchKernelGGL(id_scan_kernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, base, stride, total,
                           0LL, 0LL, 0LL, flags);
    }
}
void launch_expand_member_rows(uint8_t const* src, int const* lane_member, int cols, int S_pad, uint8_t* dst,
                               hipStream_t st) {
    long long const total = (long long)(cols) * S_pad;
    if (total > 0)
        hipLaunchKernelGGL(expand_member_rows_kernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, src,
                           lane_member, cols, S_pad, dst);
}
void launch_expand_ylane(double const* ydata, int E, int PP, int S_pad, double* ylane, int n_patch,
                         int const* patch_lane, int const* patch_entry, double const* patch_val, hipStream_t st) {
    long long const n_val = (long long)(E) * PP * 2;
    long long const total = n_val * S_pad;
    if (total > 0)
        hipLaunchKernelGGL(broadcast_y_kernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, ydata, n_val,
                           S_pad, ylane);
    long long const np = (long long)(n_patch) * PP * 2;
    if (np > 0)
        hipLaunchKernelGGL(patch_y_kernel, dim3(unsigned((np + 255) / 256)), dim3(256), 0, st, n_patch, PP * 2,
                           S_pad, patch_lane, patch_entry, patch_val, ylane);
}
void launch_converge(PairState const& ps, int it, int max_iter, double err_tol, int* active, hipStream_t st,
                     int const* it_dev) {
    if (ps.n == 0) return;
    hipLaunchKernelGGL(converge_kernel, dim3(unsigned((ps.n + 255) / 256)), dim3(256), 0, st, ps, it, it_dev,
                       max_iter, err_tol, active);
}
void launch_refine_start(PairState const& ps, DevGroup const* groups, long long const* bus_offsets, int n_groups,
                         long long total, int P, hipStream_t st) {
    if (ps.n == 0 || total == 0) return;
    hipLaunchKernelGGL(refine_reset_kernel, dim3(unsigned((ps.n + 255) / 256)), dim3(256), 0, st, ps);
    dim3 const grid(unsigned((total + 255) / 256));
    if (P == 1)
        hipLaunchKernelGGL(refine_start_kernel<1>, grid, dim3(256), 0, st, groups, bus_offsets, n_groups);
    else
        hipLaunchKernelGGL(refine_start_kernel<3>, grid, dim3(256), 0, st, groups, bus_offsets, n_groups);
}
void launch_refine_residual(PairState const& ps, DevGroup const* groups, long long const* bus_offsets, int n_groups,
                            long long total, int P, unsigned long long* gmax, hipStream_t st) {
    if (ps.n == 0 || total == 0) return;
    hipLaunchKernelGGL(refine_reset_kernel, dim3(unsigned((ps.n + 255) / 256)), dim3(256), 0, st, ps);
    dim3 const grid(unsigned((total + 255) / 256));
    if (P == 1) {
        hipLaunchKernelGGL(refine_residual_kernel<1>, grid, dim3(256), 0, st, groups, bus_offsets, n_groups);
        hipLaunchKernelGGL(refine_berr_kernel<1>, grid, dim3(256), 0, st, groups, bus_offsets, n_groups, gmax);
    } else {
        hipLaunchKernelGGL(refine_residual_kernel<3>, grid, dim3(256), 0, st, groups, bus_offsets, n_groups);
        hipLaunchKernelGGL(refine_berr_kernel<3>, grid, dim3(256), 0, st, groups, bus_offsets, n_groups, gmax);
    }
}
void launch_refine_finish(PairState const& ps, DevGroup const* groups, long long const* bus_offsets, int n_groups,
                          long long total, int P, int max_iter, double err_tol, int* active, hipStream_t st) {
    if (ps.n == 0) return;
    if (total > 0) {
        dim3 const grid(unsigned((total + 255) / 256));
        if (P == 1)
            hipLaunchKernelGGL(refine_apply_kernel<1>, grid, dim3(256), 0, st, groups, bus_offsets, n_groups);
        else
            hipLaunchKernelGGL(refine_apply_kernel<3>, grid, dim3(256), 0, st, groups, bus_offsets, n_groups);
    }
    hipLaunchKernelGGL(refine_finish_kernel, dim3(unsigned((ps.n + 255) / 256)), dim3(256), 0, st, ps, max_iter,
                       err_tol, active);
}
void launch_begin_iteration(int* it_dev, int* active, hipStream_t st) {
    hipLaunchKernelGGL(begin_iteration_kernel, dim3(1), dim3(64), 0, st, it_dev, active);
}
void launch_set_int(int* p, int v, hipStream_t st) { hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(64), 0, st, p, v); }
void launch_ic_start(int P, DevGroup const* groups, long long const* bus_offsets, int n_groups, long long total,
                     hipStream_t st) {
    if (total == 0) return;
    dim3 const grid(unsigned((total + 255) / 256));
    if (P == 1)
        hipLaunchKernelGGL(ic_start_kernel<1>, grid, dim3(256), 0, st, groups, bus_offsets, n_groups);
    else
        hipLaunchKernelGGL(ic_start_kernel<3>, grid, dim3(256), 0, st, groups, bus_offsets, n_groups);
}
void launch_retire_linear(PairState const& ps, uint8_t const* pair_linear, hipStream_t st) {
    if (ps.n == 0) return;
    hipLaunchKernelGGL(retire_linear_kernel, dim3(unsigned((ps.n + 255) / 256)), dim3(256), 0, st, ps, pair_linear);
}
void launch_init_state(PairState const& ps, hipStream_t st) {
    if (ps.n == 0) return;
    hipLaunchKernelGGL(init_state_kernel, dim3(unsigned((ps.n + 255) / 256)), dim3(256), 0, st, ps);
}
void launch_bus_results(int P, DevOutput const* outs, long long const* offsets, int n_groups, long long total,
                        hipStream_t st) {
    if (total == 0) return;
    dim3 const grid(unsigned((total + 255) / 256));
    if (P == 1)
        hipLaunchKernelGGL(bus_results_kernel<1>, grid, dim3(256), 0, st, outs, offsets, n_groups);
    else
        hipLaunchKernelGGL(bus_results_kernel<3>, grid, dim3(256), 0, st, outs, offsets, n_groups);
}
void launch_adjoint_rhs(DevGroup const* groups, DevAdjoint const* adj, long long const* bus_offsets, int n_groups,
                        long long total, hipStream_t st) {
    if (total == 0) return;
    hipLaunchKernelGGL(adjoint_rhs_kernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, groups, adj,
                       bus_offsets, n_groups);
}
void launch_adjoint_sweep(DevGroup const* groups, WorkItem const* items, int n_items, bool fp32, bool ut_sweep,
                          hipStream_t st) {
    if (n_items <= 0) return;
    dim3 const grid{unsigned(n_items)};
    if (fp32 && ut_sweep)
        hipLaunchKernelGGL((adjoint_sweep_kernel<float, true>), grid, dim3(kWave), 0, st, groups, items);
    else if (fp32)
        hipLaunchKernelGGL((adjoint_sweep_kernel<float, false>), grid, dim3(kWave), 0, st, groups, items);
    else if (ut_sweep)
        hipLaunchKernelGGL((adjoint_sweep_kernel<double, true>), grid, dim3(kWave), 0, st, groups, items);
    else
        hipLaunchKernelGGL((adjoint_sweep_kernel<double, false>), grid, dim3(kWave), 0, st, groups, items);
}
void launch_adjoint_grad(DevGroup const* groups, DevOutput const* outs, DevAdjoint const* adj,
                         long long const* offsets, int n_groups, long long total, hipStream_t st) {
    if (total == 0) return;
    hipLaunchKernelGGL(adjoint_grad_kernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, groups, outs, adj,
                       offsets, n_groups);
}
void launch_load_broadcast(double* s_inj, int8_t* lstat, double const* base_s, int8_t const* base_status, int n_load,
                           int n_lg, int P, int S, int S_pad, hipStream_t st) {
    int const rows_s = n_load * P * 2;
    long long const total = ((long long)(rows_s) + n_lg) * S_pad;
    if (total == 0) return;
    hipLaunchKernelGGL(load_broadcast_kernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0, st, s_inj, lstat,
                       base_s, base_status, rows_s, n_lg, S, S_pad);
}
void launch_load_scatter(double* s_inj, int8_t* lstat, int S, int S_pad, int const* scenario_index,
                         LoadUpdateSource const& src, LoadTable const& tab, hipStream_t st) {
    if (S == 0 || src.n_elem == 0) return;
    dim3 const grid(unsigned((S + kScatterLanes - 1) / kScatterLanes),
                    unsigned((src.n_elem + kScatterElems - 1) / kScatterElems));
    hipLaunchKernelGGL(load_scatter_kernel, grid, dim3(256), 0, st, s_inj, lstat, S, S_pad, scenario_index, src,
                       tab);
}
void launch_lowrank(DevGroup const* groups, ExtraWork const* work, int n_work, bool fp32, hipStream_t st) {
    if (n_work == 0) return;
    if (fp32)
        hipLaunchKernelGGL(lowrank_kernel<float>, dim3(unsigned(n_work)), dim3(kWave), 0, st, groups, work);
    else
        hipLaunchKernelGGL(lowrank_kernel<double>, dim3(unsigned(n_work)), dim3(kWave), 0, st, groups, work);
}
void launch_node_records(int P, DevOutput const* outs, NodeRec const* recs, long long const* tile_off, int n_groups,
                         long long tiles, void* out, ColumnSet const& cols, hipStream_t st) {
    if (tiles == 0) return;
    if (P == 1)
        hipLaunchKernelGGL((records_kernel<1, NodeRec, NodeOutput<1>>), dim3(unsigned(tiles)), dim3(256), 0, st, outs,
                           recs, tile_off, n_groups, static_cast<NodeOutput<1>*>(out), cols);
    else
        hipLaunchKernelGGL((records_kernel<3, NodeRec, NodeOutput<3>>), dim3(unsigned(tiles)), dim3(256), 0, st, outs,
                           recs, tile_off, n_groups, static_cast<NodeOutput<3>*>(out), cols);
}
void launch_branch_records(int P, DevOutput const* outs, BranchRec const* recs, long long const* tile_off,
                           int n_groups, long long tiles, void* out, ColumnSet const& cols, hipStream_t st) {
    if (tiles == 0) return;
    if (P == 1)
        hipLaunchKernelGGL((records_kernel<1, BranchRec, BranchOutput<1>>), dim3(unsigned(tiles)), dim3(256), 0, st,
                           outs, recs, tile_off, n_groups, static_cast<BranchOutput<1>*>(out), cols);
    else
        hipLaunchKernelGGL((records_kernel<3, BranchRec, BranchOutput<3>>), dim3(unsigned(tiles)), dim3(256), 0, st,
                           outs, recs, tile_off, n_groups, static_cast<BranchOutput<3>*>(out), cols);
}
void launch_appliance_records(int P, DevOutput const* outs, ApplRec const* recs, long long const* tile_off,
                              int n_groups, long long tiles, void* out, ColumnSet const& cols, hipStream_t st) {
    if (tiles == 0) return;
    if (P == 1)
        hipLaunchKernelGGL((records_kernel<1, ApplRec, ApplianceOutput<1>>), dim3(unsigned(tiles)), dim3(256), 0, st,
                           outs, recs, tile_off, n_groups, static_cast<ApplianceOutput<1>*>(out), cols);
    else
        hipLaunchKernelGGL((records_kernel<3, ApplRec, ApplianceOutput<3>>), dim3(unsigned(tiles)), dim3(256), 0, st,
                           outs, recs, tile_off, n_groups, static_cast<ApplianceOutput<3>*>(out), cols);
}

}  // namespace pgm_amd
