// se_kernels.hip -- gfx950 kernels of the batched Newton-Raphson state estimation.
//
// One thread per (bus, scenario lane); lanes are the fastest index of every per-scenario array, so a wave
// reads and writes 64 consecutive scenarios. The gain matrix of every lane is assembled row by row (a row
// only writes its own blocks and right-hand side); the Q^T blocks and Lagrange terms follow once every row
// has its Q blocks. The factorisation and the triangular solves are the block LU sweeps of the power flow
// on 4x4 blocks (launch_gain_sweep).
//
// Math restated from the reference (math_solver/newton_raphson_se_solver.hpp):
//   initial state              :219-236
//   gain matrix and rhs        :253-413 (voltage :647-679, injection :397-448, shunt :450-471,
//                              branch power :473-523, weights and products :682-775)
//   Q^T and Lagrange terms     :583-617
//   update and deviation       :777-821
//   perturbation threshold     sparse_lu_solver.hpp:352, 614-637 (block-wise off-diagonal infinity norm)
//   refinement                 sparse_lu_solver.hpp:514-612
#include "kernels.hpp"

#include <cmath>

namespace pgm_amd {
namespace dev {

namespace {
struct Cx {
    double re, im;
};
__device__ __forceinline__ Cx cx_mul(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ Cx cx_conj(Cx a) { return {a.re, -a.im}; }
__device__ __forceinline__ Cx cx_scale(Cx a, double s) { return {a.re * s, a.im * s}; }
__device__ __forceinline__ Cx cx_add(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }

// dP/dtheta, dP/dV, dQ/dtheta, dQ/dV of one measurement function against one bus
struct Jac {
    double pt, pv, qt, qv;
};
// calculate_jacobian: H, N, M, L from hm = conj(y) u_a conj(u_b) and nl = hm / |u_b|
__device__ __forceinline__ Jac jac(Cx hm, Cx nl) { return {hm.im, nl.re, -hm.re, nl.im}; }
// jacobian_diagonal_component
__device__ __forceinline__ Jac jac_diag(Cx fv, Cx f) { return {-f.im, fv.re, f.re, fv.im}; }
__device__ __forceinline__ Jac jac_add(Jac a, Jac b) { return {a.pt + b.pt, a.pv + b.pv, a.qt + b.qt, a.qv + b.qv}; }
// transpose_multiply_weight: F^T W with W = diag(1 / var_p, 1 / var_q)
__device__ __forceinline__ Jac jac_tw(Jac j, double wp, double wq) {
    return {wp * j.pt, wq * j.qt, wp * j.pv, wq * j.qv};
}
// column-major 4x4: element (r, c) at c * 4 + r
constexpr int at(int r, int c) { return c * 4 + r; }
// multiply_add_jacobian_blocks_lhs: G(0..1, 0..1) += F . X
__device__ __forceinline__ void add_lhs(double (&b)[16], Jac f, Jac x) {
    b[at(0, 0)] += f.pt * x.pt + f.pv * x.qt;
    b[at(0, 1)] += f.pt * x.pv + f.pv * x.qv;
    b[at(1, 0)] += f.qt * x.pt + f.qv * x.qt;
    b[at(1, 1)] += f.qt * x.pv + f.qv * x.qv;
}
// multiply_add_jacobian_blocks_rhs: eta += F . (z - f(x))
__device__ __forceinline__ void add_rhs(double (&r)[4], Jac f, double zp, double zq, Cx fx) {
    double const dp = zp - fx.re, dq = zq - fx.im;
    r[0] += f.pt * dp + f.pv * dq;
    r[1] += f.qt * dp + f.qv * dq;
}
// add_injection_jacobian: Q block (rows 2..3, columns 0..1)
__device__ __forceinline__ void add_q(double (&b)[16], Jac j) {
    b[at(2, 0)] += j.pt;
    b[at(2, 1)] += j.pv;
    b[at(3, 0)] += j.qt;
    b[at(3, 1)] += j.qv;
}

__device__ __forceinline__ double bits_to_double(unsigned long long b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ unsigned long long double_to_bits(double d) {
    return static_cast<unsigned long long>(__double_as_longlong(d));
}
}  // namespace

// one flat index over (bus, lane) with the lane fastest
__device__ __forceinline__ bool bus_lane(SeDev const& g, int& p, int& s) {
    long long const i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)g.n * g.S_pad) return false;
    p = int(i / g.S_pad);
    s = int(i - (long long)p * g.S_pad);
    return true;
}

__global__ void se_init_kernel(SeDev g, double const* __restrict__ mean_shift) {
    int p, s;
    if (!bus_lane(g, p, s)) return;
    size_t const sp = size_t(g.S_pad);
    int const mm = g.bus_mm[p];
    size_t const pair = size_t(mm) * sp + s;
    double theta = mean_shift[pair] + g.phase[p];
    double v = 1.0;
    int const vf = g.v_flag[size_t(p) * sp + s];
    if (vf) {
        double const re = g.v_val[(size_t(p) * 3) * sp + s], im = g.v_val[(size_t(p) * 3 + 1) * sp + s];
        if (vf == 2) {
            theta = atan2(im, re);
            v = sqrt(re * re + im * im);
        } else {
            v = re;
        }
    }
    size_t const b = size_t(p) * 4;
    g.x[(b + 0) * sp + s] = theta;
    g.x[(b + 1) * sp + s] = v;
    g.x[(b + 2) * sp + s] = 0.0;
    g.x[(b + 3) * sp + s] = 0.0;
    g.u[(size_t(p) * 2) * sp + s] = v * cos(theta);
    g.u[(size_t(p) * 2 + 1) * sp + s] = v * sin(theta);
}

__global__ void se_pair_reset_kernel(SeDev g) {
    long long const i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)g.n_mm * g.S_pad) return;
    g.norm[i] = 0ull;
    g.pert[i] = 0;
}

// the measurements of one row: the diagonal block, the right-hand side, and the off-diagonal blocks of the
// row (prepare_matrix_and_rhs)
__global__ void se_gain_kernel(SeDev g) {
    int p, s;
    if (!bus_lane(g, p, s)) return;
    size_t const sp = size_t(g.S_pad);
    int const mm = g.bus_mm[p];
    size_t const pair = size_t(mm) * sp + s;
    if (g.done[pair] != ACTIVE) return;
    auto X = [&](int bus, int c) { return g.x[(size_t(bus) * 4 + c) * sp + s]; };
    auto U = [&](int bus) { return Cx{g.u[(size_t(bus) * 2) * sp + s], g.u[(size_t(bus) * 2 + 1) * sp + s]}; };
    Cx const ui = U(p);
    double const inv_i = 1.0 / X(p, 1);
    Cx const ui_ui = cx_mul(ui, cx_conj(ui));
    double D[16], R[4];
#pragma unroll
    for (int q = 0; q < 16; ++q) D[q] = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) R[q] = 0.0;
    int const vf = g.v_flag[size_t(p) * sp + s];
    bool const inj = g.inj_flag[size_t(p) * sp + s] != 0;
    int const dg = g.diag[p];
    for (int k = g.row_ptr[p]; k < g.row_ptr[p + 1]; ++k) {
        int const c = g.col[k];
        bool const is_diag = k == dg;
        double B[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) B[q] = 0.0;
        if (is_diag && vf) {
            // process_voltage_measurements
            double const re = g.v_val[(size_t(p) * 3) * sp + s], im = g.v_val[(size_t(p) * 3 + 1) * sp + s];
            double const w_v = 1.0 / g.v_val[(size_t(p) * 3 + 2) * sp + s];
            double const abs_v = vf == 2 ? sqrt(re * re + im * im) : re;
            double const delta_v = abs_v - X(p, 1);
            double w_t = 0.0, delta_t = 0.0;
            if (vf == 2) {
                delta_t = atan2(im, re) - X(p, 0);
                w_t = w_v;
            } else if (p == g.vref[pair] && !g.has_angle[pair]) {
                delta_t = 0.0 - X(p, 0);
                w_t = w_v;
            }
            D[at(0, 0)] += w_t;
            D[at(1, 1)] += w_v;
            R[0] += w_t * delta_t;
            R[1] += w_v * delta_v;
        }
        if (g.fill[k]) {
            if (!is_diag) {
                for (int q = 0; q < 16; ++q) g.raw_a[(size_t(k) * 16 + q) * sp + s] = 0.0;
            }
            continue;
        }
        double (&blk)[16] = is_diag ? D : B;
        Cx const uj = U(c);
        double const inv_j = 1.0 / X(c, 1);
        Cx const ui_uj = cx_mul(ui, cx_conj(uj));
        Cx const uj_ui = cx_mul(uj, cx_conj(ui));
        Cx const uj_uj = cx_mul(uj, cx_conj(uj));
        // power measured at one branch side (process_branch_power_measurement): chi is the measured side's
        // bus, psi the other; row_major when chi is this row
        auto branch_power = [&](Cx y_xx, Cx y_xm, bool row_major, double const* val, size_t idx) {
            double const zp = val[(idx * 4) * sp + s], zq = val[(idx * 4 + 1) * sp + s];
            double const wp = 1.0 / val[(idx * 4 + 2) * sp + s], wq = 1.0 / val[(idx * 4 + 3) * sp + s];
            Cx const chi_chi = row_major ? ui_ui : uj_uj;
            Cx const chi_psi = row_major ? ui_uj : uj_ui;
            double const inv_chi = row_major ? inv_i : inv_j, inv_psi = row_major ? inv_j : inv_i;
            Cx const hm1 = cx_mul(cx_conj(y_xx), chi_chi);
            Cx const nl1 = cx_scale(hm1, inv_chi);
            Cx const hm2 = cx_mul(cx_conj(y_xm), chi_psi);
            Cx const nl2 = cx_scale(hm2, inv_psi);
            Cx const f = cx_add(hm1, hm2);
            Cx const fv = cx_scale(f, inv_chi);
            Jac const j_cc = jac_add(jac(hm1, nl1), jac_diag(fv, f));
            Jac const j_cp = jac(hm2, nl2);
            Jac const left = row_major ? j_cc : j_cp;
            Jac const right = row_major ? j_cp : j_cc;
            Jac const ftw = jac_tw(left, wp, wq);
            add_lhs(D, ftw, left);
            add_lhs(blk, ftw, right);
            add_rhs(R, ftw, zp, zq, f);
        };
        for (int e = g.el_ptr[k]; e < g.el_ptr[k + 1]; ++e) {
            int const code = g.el[e];
            int const kind = code & 3, idx = code >> 2;
            if (kind == 0) {
                // process_shunt_measurement (diagonal entries only)
                if (!g.sh_flag[size_t(idx) * sp + s]) continue;
                Cx const ys = {g.ysh[2 * idx], g.ysh[2 * idx + 1]};
                double const zp = g.sh_val[(size_t(idx) * 4) * sp + s], zq = g.sh_val[(size_t(idx) * 4 + 1) * sp + s];
                double const wp = 1.0 / g.sh_val[(size_t(idx) * 4 + 2) * sp + s];
                double const wq = 1.0 / g.sh_val[(size_t(idx) * 4 + 3) * sp + s];
                Cx const hm = cx_mul(cx_conj({-ys.re, -ys.im}), ui_ui);
                Cx const nl = cx_scale(hm, inv_i);
                Jac const jb = jac_add(jac(hm, nl), jac_diag(nl, hm));
                Jac const ftw = jac_tw(jb, wp, wq);
                add_lhs(D, ftw, jb);
                add_rhs(R, ftw, zp, zq, hm);
                continue;
            }
            double const* y = g.ybr + size_t(idx) * 8;
            Cx const yff = {y[0], y[1]}, yft = {y[2], y[3]}, ytf = {y[4], y[5]}, ytt = {y[6], y[7]};
            bool const bft = kind == 1;  // this row is the branch's from side
            if (g.bf_flag[size_t(idx) * sp + s]) branch_power(yff, yft, bft, g.bf_val, size_t(idx));
            if (g.bt_flag[size_t(idx) * sp + s]) branch_power(ytt, ytf, !bft, g.bt_val, size_t(idx));
        }
        if (inj) {
            // process_injection_row, and process_injection_diagonal on the diagonal entry
            Cx const yij = {g.yent[2 * size_t(k)], g.yent[2 * size_t(k) + 1]};
            Cx const hm = cx_mul(cx_conj(yij), ui_uj);
            Cx const nl = cx_scale(hm, inv_j);
            add_q(blk, jac(hm, nl));
            Cx const fv = cx_scale(hm, inv_i);
            add_q(D, jac_diag(fv, hm));
            R[2] -= hm.re;
            R[3] -= hm.im;
            if (is_diag) {
                R[2] += g.inj_val[(size_t(p) * 4) * sp + s];
                R[3] += g.inj_val[(size_t(p) * 4 + 1) * sp + s];
                D[at(2, 2)] = -g.inj_val[(size_t(p) * 4 + 2) * sp + s];
                D[at(3, 3)] = -g.inj_val[(size_t(p) * 4 + 3) * sp + s];
            }
        } else if (is_diag) {
            // virtually_remove_constraints
            D[at(2, 2)] = -1.0;
            D[at(3, 3)] = -1.0;
        }
        if (!is_diag) {
            for (int q = 0; q < 16; ++q) g.raw_a[(size_t(k) * 16 + q) * sp + s] = B[q];
        }
    }
    for (int q = 0; q < 16; ++q) g.raw_a[(size_t(dg) * 16 + q) * sp + s] = D[q];
    for (int q = 0; q < 4; ++q) g.raw_b[(size_t(p) * 4 + q) * sp + s] = R[q];
}

// Q^T(row, col) = Q(col, row)^T, eta(row) += Q^T(row, col) phi(col) over the row's Y-bus entries, and the
// row's sum of the off-diagonal blocks' infinity norms into the pair's matrix norm
__global__ void se_qt_kernel(SeDev g) {
    int p, s;
    if (!bus_lane(g, p, s)) return;
    size_t const sp = size_t(g.S_pad);
    size_t const pair = size_t(g.bus_mm[p]) * sp + s;
    if (g.done[pair] != ACTIVE) return;
    auto A = [&](int k, int q) -> double& { return g.raw_a[(size_t(k) * 16 + q) * sp + s]; };
    double e0 = 0.0, e1 = 0.0;
    int const dg = g.diag[p];
    for (int k = g.row_ptr[p]; k < g.row_ptr[p + 1]; ++k) {
        if (g.fill[k]) continue;
        int const c = g.col[k];
        int const t = g.transpose[k];
        double const q20 = A(t, at(2, 0)), q21 = A(t, at(2, 1)), q30 = A(t, at(3, 0)), q31 = A(t, at(3, 1));
        A(k, at(0, 2)) = q20;
        A(k, at(0, 3)) = q30;
        A(k, at(1, 2)) = q21;
        A(k, at(1, 3)) = q31;
        double const php = g.x[(size_t(c) * 4 + 2) * sp + s], phq = g.x[(size_t(c) * 4 + 3) * sp + s];
        e0 += q20 * php + q30 * phq;
        e1 += q21 * php + q31 * phq;
    }
    g.raw_b[(size_t(p) * 4) * sp + s] += e0;
    g.raw_b[(size_t(p) * 4 + 1) * sp + s] += e1;
    double row_norm = 0.0;
    for (int k = g.row_ptr[p]; k < g.row_ptr[p + 1]; ++k) {
        if (k == dg) continue;
        double bn = 0.0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            double rs = 0.0;
#pragma unroll
            for (int c = 0; c < 4; ++c) rs += fabs(A(k, at(r, c)));
            bn = rs > bn ? rs : bn;
        }
        row_norm += bn;
    }
    atomicMax(&g.norm[pair], double_to_bits(row_norm));
}

// perturbation threshold epsilon_perturbation * norm for the pairs that use it
__global__ void se_threshold_kernel(SeDev g) {
    long long const i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)g.n_mm * g.S_pad) return;
    g.pthr[i] = g.use_pert[i] ? 1e-13 * bits_to_double(g.norm[i]) : 0.0;
}

// refinement start: the factor sweep's solution is the first iterate
__global__ void se_refine_begin_kernel(SeDev g) {
    int p, s;
    if (!bus_lane(g, p, s)) return;
    size_t const sp = size_t(g.S_pad);
    size_t const pair = size_t(g.bus_mm[p]) * sp + s;
    if (g.done[pair] != REFINE) return;
#pragma unroll
    for (int r = 0; r < 4; ++r) g.xacc[(size_t(p) * 4 + r) * sp + s] = g.y[(size_t(p) * 4 + r) * sp + s];
}

// largest denominator and backward error of the pairs in REFINE start at zero before every residual
__global__ void se_refine_reset_kernel(SeDev g) {
    long long const i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)g.n_mm * g.S_pad) return;
    g.rmax[i] = 0ull;
    g.berr[i] = 0ull;
}

// r = b - A x and |b| + |A| |x| per row (calculate_residual, iterate_and_backward_error's denominator)
__global__ void se_residual_kernel(SeDev g) {
    int p, s;
    if (!bus_lane(g, p, s)) return;
    size_t const sp = size_t(g.S_pad);
    size_t const pair = size_t(g.bus_mm[p]) * sp + s;
    if (g.done[pair] != REFINE) return;
    double r[4], d[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double const b = g.raw_b[(size_t(p) * 4 + q) * sp + s];
        r[q] = b;
        d[q] = fabs(b);
    }
    for (int k = g.row_ptr[p]; k < g.row_ptr[p + 1]; ++k) {
        int const c = g.col[k];
        double xc[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) xc[q] = g.xacc[(size_t(c) * 4 + q) * sp + s];
#pragma unroll
        for (int rr = 0; rr < 4; ++rr)
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) {
                double const a = g.raw_a[(size_t(k) * 16 + at(rr, cc)) * sp + s];
                r[rr] -= a * xc[cc];
                d[rr] += fabs(a) * fabs(xc[cc]);
            }
    }
    double dm = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        g.y[(size_t(p) * 4 + q) * sp + s] = r[q];
        g.rden[(size_t(p) * 4 + q) * sp + s] = d[q];
        dm = d[q] > dm ? d[q] : dm;
    }
    atomicMax(&g.rmax[pair], double_to_bits(dm));
}

// backward error max |r| / max(den, 1e-4 max den) per pair
__global__ void se_berr_kernel(SeDev g) {
    int p, s;
    if (!bus_lane(g, p, s)) return;
    size_t const sp = size_t(g.S_pad);
    size_t const pair = size_t(g.bus_mm[p]) * sp + s;
    if (g.done[pair] != REFINE) return;
    double const cap = 1e-4 * bits_to_double(g.rmax[pair]);
    double be = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double const num = fabs(g.y[(size_t(p) * 4 + q) * sp + s]);
        double const den = fmax(g.rden[(size_t(p) * 4 + q) * sp + s], cap);
        double const e = den > 0.0 ? num / den : 0.0;
        be = e > be ? e : be;
    }
    atomicMax(&g.berr[pair], double_to_bits(be));
}

// angle offset of the slack bus when no voltage angle is measured (iterate_unknown :780-791)
__global__ void se_offset_kernel(SeDev g) {
    long long const i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)g.n_mm * g.S_pad) return;
    if (g.done[i] != ACTIVE) return;
    int const mm = int(i / g.S_pad);
    size_t const s = size_t(i - (long long)mm * g.S_pad), sp = size_t(g.S_pad);
    if (g.has_angle[i]) {
        g.offset[i] = 0.0;
        return;
    }
    int const sl = g.slack[mm];
    double const* dx = g.pert[i] ? g.xacc : g.y;
    g.offset[i] = g.x[(size_t(sl) * 4) * sp + s] + dx[(size_t(sl) * 4) * sp + s];
}

__global__ void se_update_kernel(SeDev g) {
    int p, s;
    if (!bus_lane(g, p, s)) return;
    size_t const sp = size_t(g.S_pad);
    size_t const pair = size_t(g.bus_mm[p]) * sp + s;
    if (g.done[pair] != ACTIVE) return;
    double const* dx = g.pert[pair] ? g.xacc : g.y;
    size_t const b = size_t(p) * 4;
    double const th = g.x[(b + 0) * sp + s] + dx[(b + 0) * sp + s] - g.offset[pair];
    double const v = g.x[(b + 1) * sp + s] + dx[(b + 1) * sp + s];
    g.x[(b + 0) * sp + s] = th;
    g.x[(b + 1) * sp + s] = v;
    if (g.inj_flag[size_t(p) * sp + s]) {
        if (g.inj_val[(size_t(p) * 4 + 2) * sp + s] == 0.0) g.x[(b + 2) * sp + s] += dx[(b + 2) * sp + s];
        if (g.inj_val[(size_t(p) * 4 + 3) * sp + s] == 0.0) g.x[(b + 3) * sp + s] += dx[(b + 3) * sp + s];
    }
    double const ur = v * cos(th), ui = v * sin(th);
    double const dr = ur - g.u[(size_t(p) * 2) * sp + s], di = ui - g.u[(size_t(p) * 2 + 1) * sp + s];
    g.u[(size_t(p) * 2) * sp + s] = ur;
    g.u[(size_t(p) * 2 + 1) * sp + s] = ui;
    atomicMax(&g.devmax[pair], double_to_bits(sqrt(dr * dr + di * di)));
}

}  // namespace dev

namespace {
dim3 bus_grid(SeDev const& g) { return dim3(unsigned(((long long)g.n * g.S_pad + 255) / 256)); }
dim3 pair_grid(SeDev const& g) { return dim3(unsigned(((long long)g.n_mm * g.S_pad + 255) / 256)); }
}  // namespace

void launch_se_init(SeDev const& g, double const* mean_shift, hipStream_t st) {
    if (g.n == 0) return;
    hipLaunchKernelGGL(dev::se_init_kernel, bus_grid(g), dim3(256), 0, st, g, mean_shift);
}
void launch_se_pair_reset(SeDev const& g, hipStream_t st) {
    if (g.n_mm == 0) return;
    hipLaunchKernelGGL(dev::se_pair_reset_kernel, pair_grid(g), dim3(256), 0, st, g);
}
void launch_se_gain(SeDev const& g, hipStream_t st) {
    if (g.n == 0) return;
    hipLaunchKernelGGL(dev::se_gain_kernel, bus_grid(g), dim3(256), 0, st, g);
    hipLaunchKernelGGL(dev::se_qt_kernel, bus_grid(g), dim3(256), 0, st, g);
    hipLaunchKernelGGL(dev::se_threshold_kernel, pair_grid(g), dim3(256), 0, st, g);
}
void launch_se_refine_begin(SeDev const& g, hipStream_t st) {
    if (g.n == 0) return;
    hipLaunchKernelGGL(dev::se_refine_begin_kernel, bus_grid(g), dim3(256), 0, st, g);
}
void launch_se_residual(SeDev const& g, hipStream_t st) {
    if (g.n == 0) return;
    hipLaunchKernelGGL(dev::se_refine_reset_kernel, pair_grid(g), dim3(256), 0, st, g);
    hipLaunchKernelGGL(dev::se_residual_kernel, bus_grid(g), dim3(256), 0, st, g);
}
void launch_se_berr(SeDev const& g, hipStream_t st) {
    if (g.n == 0) return;
    hipLaunchKernelGGL(dev::se_berr_kernel, bus_grid(g), dim3(256), 0, st, g);
}
void launch_se_update(SeDev const& g, hipStream_t st) {
    if (g.n == 0) return;
    hipLaunchKernelGGL(dev::se_offset_kernel, pair_grid(g), dim3(256), 0, st, g);
    hipLaunchKernelGGL(dev::se_update_kernel, bus_grid(g), dim3(256), 0, st, g);
}

}  // namespace pgm_amd
