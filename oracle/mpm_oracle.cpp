// mpm_oracle.cpp — FP64 CPU restatement of the MLS-MPM substep.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg load it, as the checker or the timed CPU baseline.  The engine never
// links it and has no CPU fallback.
//
// The reference solver (libs/cuda-engine/source/MPM.cu) is CUDA-only and Windows-only, so it cannot be
// compiled here; this is a hand transcription of its substep semantics in double precision:
//   grid reset          MPM.cu:46-51
//   P2G                 MPM.cu:53-129  (Tait pressure :68-74, Q :75-80, weights :92-110, scatter :112-128)
//   grid update         MPM.cu:131-177 (normalise :151-154, walls :159-168, empty node :169-172)
//   G2P                 MPM.cu:196-266 (gather :227-251, advect :253-255, F/J :258-265)
//   out-of-bounds       MPM.cu:267-301
//   det2D / matMul2D    cuda-internal.cuh:44-56
// with the reference's compile-time constants (constants.h:72-88) turned into MpmParams fields.
//
// Materials the reference does not have (its only stress is the Tait fluid) are written from the
// literature and used identically by the engine:
//   jelly: fixed-corotated  tau = 2 mu (F - R) F^T + lambda (J - 1) J I,  J = det F,  R = polar rotation
//          (Stomakhin, Howes, Schroeder, Teran 2012, "Energetically consistent invertible elasticity").
//   snow:  same energy with mu, lambda scaled by exp(hardening (1 - Jp)); after F <- A F the singular
//          values are clamped to [1 - theta_c, 1 + theta_s] and Jp absorbs the removed volume
//          (Stomakhin, Schroeder, Chai, Teran, Selle 2013, "A material point method for snow simulation").
// The Kirchhoff stress enters P2G as Q = m C - dt V0 4 inv_dx^2 tau, which for the Tait fluid
// (tau = -J p I) is exactly the reference's Q = m C + coeff sigma with coeff = -dt V0 J 4 inv_dx^2 (:71-80).
//
// Besides the serial loop (deterministic), an OpenMP loop with per-thread private grids reduced in fixed
// thread order is provided (deterministic for a given thread count); it is the timed CPU baseline.
//
// Binning restatement: FP32 bin keys and in-bin stencil nodes computed exactly as the engine computes them
// (fmaf(x, inv_dx, halo), the contracted form of gridCoord, MPM.cu:84), std::stable_sort, and the layered
// order inside bins (oracle_layered_order) — the bit-exact order the engine's binning must reproduce.
//
// MLP restatement: forward, backward and Adam of the 4-64-64-3 SiLU closure in FP64, for gradient checks.
//
// Frame restatement: the reference renderer's background and particle blend, serial (oracle_render).
#include "../include/mpm_engine.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct P64 {
    int W, H, halo, wall;
    double dx, inv_dx, sim_w, sim_h, dt, gx, gy, mass, vol;
    double tait_k, tait_gamma, p_min, p_max, mu, lambda, theta_c, theta_s, hardening, j_min, j_max;
};

P64 promote(const MpmParams& p)
{
    P64 q;
    q.W = p.grid_w;
    q.H = p.grid_h;
    q.halo = p.halo;
    q.wall = p.wall;
    q.dx = p.dx;
    q.inv_dx = p.inv_dx;
    q.sim_w = p.sim_w;
    q.sim_h = p.sim_h;
    q.dt = p.dt;
    q.gx = p.gravity_x;
    q.gy = p.gravity_y;
    q.mass = p.mass;
    q.vol = p.vol;
    q.tait_k = p.tait_k;
    q.tait_gamma = p.tait_gamma;
    q.p_min = p.p_min;
    q.p_max = p.p_max;
    q.mu = p.mu;
    q.lambda = p.lambda;
    q.theta_c = p.theta_c;
    q.theta_s = p.theta_s;
    q.hardening = p.hardening;
    q.j_min = p.j_min;
    q.j_max = p.j_max;
    return q;
}

// Quadratic B-spline weights of the three stencil nodes, MPM.cu:92-100.
inline void bspline(double x, const P64& p, int& base, double& f, double w[3])
{
    const double g = x * p.inv_dx + (double)p.halo;  // gridCoord, MPM.cu:84
    base = (int)std::floor(g - 0.5);                  // MPM.cu:94
    f = g - (double)base;                             // MPM.cu:95
    w[0] = 0.5 * (1.5 - f) * (1.5 - f);
    w[1] = 0.75 - (f - 1.0) * (f - 1.0);
    w[2] = 0.5 * (f - 0.5) * (f - 0.5);
}

// Rotation R maximising tr(R^T F) over SO(2): the rotation factor of the polar decomposition F = R S
// (angle atan2(F10 - F01, F00 + F11)).  R = I when F00 + F11 = F10 - F01 = 0.
inline void polar_rotation(const double F[4], double R[4])
{
    const double x = F[0] + F[3];
    const double y = F[2] - F[1];
    const double r = std::hypot(x, y);
    double c = 1.0, s = 0.0;
    if (r > 0.0) {
        c = x / r;
        s = y / r;
    }
    R[0] = c;
    R[1] = -s;
    R[2] = s;
    R[3] = c;
}

// Signed 2x2 SVD F = U diag(sig) V^T with U, V rotations: polar F = R S, then the symmetric
// eigen-decomposition S = V diag(sig) V^T by one Jacobi rotation; U = R V.  sig[1] < 0 iff det F < 0.
inline void svd2(const double F[4], double U[4], double sig[2], double V[4])
{
    double R[4];
    polar_rotation(F, R);
    // S = R^T F
    const double S00 = R[0] * F[0] + R[2] * F[2];
    const double S01 = R[0] * F[1] + R[2] * F[3];
    const double S11 = R[1] * F[1] + R[3] * F[3];
    // Symmetrise (exact in exact arithmetic).
    const double S10 = R[1] * F[0] + R[3] * F[2];
    const double off = 0.5 * (S01 + S10);
    double c = 1.0, s = 0.0;
    if (off != 0.0) {
        const double tau = 0.5 * (S11 - S00) / off;
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        c = 1.0 / std::sqrt(1.0 + t * t);
        s = t * c;
    }
    // V = [[c, s], [-s, c]] diagonalises S: V^T S V = diag.
    V[0] = c;
    V[1] = s;
    V[2] = -s;
    V[3] = c;
    sig[0] = c * c * S00 - 2.0 * c * s * off + s * s * S11;
    sig[1] = s * s * S00 + 2.0 * c * s * off + c * c * S11;
    U[0] = R[0] * V[0] + R[1] * V[2];
    U[1] = R[0] * V[1] + R[1] * V[3];
    U[2] = R[2] * V[0] + R[3] * V[2];
    U[3] = R[2] * V[1] + R[3] * V[3];
}

// Fixed-corotated Kirchhoff stress tau = 2 mu (F - R) F^T + lambda (J - 1) J I, J = det F.
inline void corotated_tau(const double F[4], double mu, double lambda, double tau[4])
{
    double R[4];
    polar_rotation(F, R);
    const double J = F[0] * F[3] - F[1] * F[2];
    const double D[4] = {F[0] - R[0], F[1] - R[1], F[2] - R[2], F[3] - R[3]};
    // (F - R) F^T
    const double M00 = D[0] * F[0] + D[1] * F[1];
    const double M01 = D[0] * F[2] + D[1] * F[3];
    const double M10 = D[2] * F[0] + D[3] * F[1];
    const double M11 = D[2] * F[2] + D[3] * F[3];
    const double vol = lambda * (J - 1.0) * J;
    tau[0] = 2.0 * mu * M00 + vol;
    tau[1] = 2.0 * mu * M01;
    tau[2] = 2.0 * mu * M10;
    tau[3] = 2.0 * mu * M11 + vol;
}

// Kirchhoff stress of one particle.
inline void particle_tau(const P64& p, int mat, const double F[4], double J, double tau[4])
{
    if (mat == MPM_MAT_FLUID) {
        // Tait pressure on the tracked J, MPM.cu:68-69; tau = J sigma with sigma = -p I (:71).
        const double p_hat = p.tait_k * (std::pow(1.0 / J, p.tait_gamma) - 1.0);
        const double pr = std::fmin(std::fmax(p_hat, p.p_min), p.p_max);  // fminf/fmaxf: NaN-ignoring
        tau[0] = -J * pr;
        tau[1] = 0.0;
        tau[2] = 0.0;
        tau[3] = -J * pr;
    } else if (mat == MPM_MAT_JELLY) {
        corotated_tau(F, p.mu, p.lambda, tau);
    } else {
        // Snow: J slot holds the plastic volume ratio Jp.
        const double h = std::exp(p.hardening * (1.0 - J));
        corotated_tau(F, p.mu * h, p.lambda * h, tau);
    }
}

struct State {
    int n;
    double* x;
    double* v;
    double* C;
    double* F;
    double* J;
    const uint8_t* mat;
};

// P2G of particles [lo, hi) into grid (3 doubles per node: mom_x, mom_y, mass), MPM.cu:53-129.
void p2g_range(const P64& p, const State& s, int lo, int hi, double* grid)
{
    const double D = p.dt * p.vol * 4.0 * p.inv_dx * p.inv_dx;
    for (int i = lo; i < hi; ++i) {
        double tau[4];
        particle_tau(p, s.mat[i], &s.F[4 * i], s.J[i], tau);
        const double* C = &s.C[4 * i];
        const double Q[4] = {p.mass * C[0] - D * tau[0], p.mass * C[1] - D * tau[1], p.mass * C[2] - D * tau[2],
                             p.mass * C[3] - D * tau[3]};
        int bx, by;
        double fx, fy, wx[3], wy[3];
        bspline(s.x[2 * i], p, bx, fx, wx);
        bspline(s.x[2 * i + 1], p, by, fy, wy);
        // Memory-safety clamp of the stencil base (never active for positions inside the domain).
        bx = std::min(std::max(bx, 0), p.W - 3);
        by = std::min(std::max(by, 0), p.H - 3);
        const double vx = s.v[2 * i], vy = s.v[2 * i + 1];
        for (int a = 0; a < 3; ++a) {
            for (int b = 0; b < 3; ++b) {
                const double w = wx[a] * wy[b];
                const double dpx = ((double)a - fx) * p.dx;
                const double dpy = ((double)b - fy) * p.dx;
                const double afx = dpx * Q[0] + dpy * Q[1];  // MPM.cu:121
                const double afy = dpx * Q[2] + dpy * Q[3];
                const size_t node = (size_t)(by + b) * (size_t)p.W + (size_t)(bx + a);  // MPM.cu:122
                grid[3 * node + 0] += w * (vx * p.mass + afx);  // MPM.cu:124-126
                grid[3 * node + 1] += w * (vy * p.mass + afy);
                grid[3 * node + 2] += w * p.mass;
            }
        }
    }
}

// Grid update, MPM.cu:131-177: normalise, gravity, outward-only walls, empty nodes to zero.
// grid holds (mom_x, mom_y, mass) and is rewritten to (v_x, v_y, mass).
void grid_update_range(const P64& p, double* grid, size_t lo, size_t hi)
{
    for (size_t idx = lo; idx < hi; ++idx) {
        double* g = &grid[3 * idx];
        if (g[2] > 0.0) {
            const double inv_m = 1.0 / g[2];
            double vx = g[0] * inv_m + p.gx * p.dt;
            double vy = g[1] * inv_m + p.gy * p.dt;
            const int x = (int)(idx % (size_t)p.W);
            const int y = (int)(idx / (size_t)p.W);
            if (x <= p.wall && vx < 0.0) {
                vx = 0.0;
            } else if (x >= p.W - 1 - p.wall && vx > 0.0) {
                vx = 0.0;
            }
            if (y <= p.wall && vy < 0.0) {
                vy = 0.0;
            } else if (y >= p.H - 1 - p.wall && vy > 0.0) {
                vy = 0.0;
            }
            g[0] = vx;
            g[1] = vy;
        } else {
            g[0] = 0.0;
            g[1] = 0.0;
        }
    }
}

// handleOOB, MPM.cu:267-301.
inline void handle_oob(const P64& p, double& px, double& py, double& vx, double& vy)
{
    if (!std::isfinite(px) || !std::isfinite(vx)) {
        px = 0.5 * p.sim_w;
        vx = 0.0;
    }
    if (!std::isfinite(py) || !std::isfinite(vy)) {
        py = 0.5 * p.sim_h;
        vy = 0.0;
    }
    if (px < 0.0) {
        px = 0.0;
        vx = std::fmax(vx, 0.0);
    }
    if (px > p.sim_w) {
        px = p.sim_w;
        vx = std::fmin(vx, 0.0);
    }
    if (py < 0.0) {
        py = 0.0;
        vy = std::fmax(vy, 0.0);
    }
    if (py > p.sim_h) {
        py = p.sim_h;
        vy = std::fmin(vy, 0.0);
    }
}

// Snow return mapping on F (in place), Jp updated (Stomakhin et al. 2013, section 7).
inline void snow_project(const P64& p, double F[4], double& Jp)
{
    double U[4], sig[2], V[4];
    svd2(F, U, sig, V);
    const double lo = 1.0 - p.theta_c, hi = 1.0 + p.theta_s;
    const double s0 = std::min(std::max(sig[0], lo), hi);
    const double s1 = std::min(std::max(sig[1], lo), hi);
    Jp *= (sig[0] * sig[1]) / (s0 * s1);
    // F = U diag(s) V^T
    F[0] = U[0] * s0 * V[0] + U[1] * s1 * V[1];
    F[1] = U[0] * s0 * V[2] + U[1] * s1 * V[3];
    F[2] = U[2] * s0 * V[0] + U[3] * s1 * V[1];
    F[3] = U[2] * s0 * V[2] + U[3] * s1 * V[3];
}

// G2P of particles [lo, hi) from grid (v_x, v_y, mass per node), MPM.cu:196-266.
void g2p_range(const P64& p, const State& s, int lo, int hi, const double* grid)
{
    const double sB = 4.0 * p.inv_dx * p.inv_dx;  // MPM.cu:252
    for (int i = lo; i < hi; ++i) {
        int bx, by;
        double fx, fy, wx[3], wy[3];
        bspline(s.x[2 * i], p, bx, fx, wx);
        bspline(s.x[2 * i + 1], p, by, fy, wy);
        bx = std::min(std::max(bx, 0), p.W - 3);
        by = std::min(std::max(by, 0), p.H - 3);
        double vx = 0.0, vy = 0.0, B[4] = {0.0, 0.0, 0.0, 0.0};
        for (int a = 0; a < 3; ++a) {
            for (int b = 0; b < 3; ++b) {
                const double w = wx[a] * wy[b];
                const double dpx = ((double)a - fx) * p.dx;
                const double dpy = ((double)b - fy) * p.dx;
                const size_t node = (size_t)(by + b) * (size_t)p.W + (size_t)(bx + a);
                const double gvx = grid[3 * node + 0], gvy = grid[3 * node + 1];
                vx += w * gvx;
                vy += w * gvy;
                B[0] += w * gvx * dpx;  // MPM.cu:246-249
                B[1] += w * gvx * dpy;
                B[2] += w * gvy * dpx;
                B[3] += w * gvy * dpy;
            }
        }
        double px = s.x[2 * i] + vx * p.dt;  // MPM.cu:253-254
        double py = s.x[2 * i + 1] + vy * p.dt;
        double* C = &s.C[4 * i];
        C[0] = sB * B[0];  // MPM.cu:255 (before the OOB handling)
        C[1] = sB * B[1];
        C[2] = sB * B[2];
        C[3] = sB * B[3];
        handle_oob(p, px, py, vx, vy);  // MPM.cu:256
        s.x[2 * i] = px;
        s.x[2 * i + 1] = py;
        s.v[2 * i] = vx;
        s.v[2 * i + 1] = vy;
        const double A[4] = {1.0 + C[0] * p.dt, C[1] * p.dt, C[2] * p.dt, 1.0 + C[3] * p.dt};  // MPM.cu:258-262
        double* F = &s.F[4 * i];
        const double F0[4] = {F[0], F[1], F[2], F[3]};
        F[0] = A[0] * F0[0] + A[1] * F0[2];  // matMul2D, cuda-internal.cuh:45-56
        F[1] = A[0] * F0[1] + A[1] * F0[3];
        F[2] = A[2] * F0[0] + A[3] * F0[2];
        F[3] = A[2] * F0[1] + A[3] * F0[3];
        if (s.mat[i] == MPM_MAT_SNOW) {
            snow_project(p, F, s.J[i]);
        } else {
            const double detA = A[0] * A[3] - A[1] * A[2];  // det2D, cuda-internal.cuh:44
            s.J[i] = std::fmin(std::fmax(s.J[i] * detA, p.j_min), p.j_max);  // MPM.cu:264-265
        }
    }
}

int run(const MpmParams* params, int32_t n, double* x, double* v, double* C, double* F, double* J,
        const uint8_t* mat, int32_t substeps, int32_t threads, double* grid_out)
{
    if (!params || n < 0 || substeps < 0 || (n > 0 && (!x || !v || !C || !F || !J || !mat))) { return -1; }
    const P64 p = promote(*params);
    if (p.W < 3 || p.H < 3) { return -1; }
    const size_t G = (size_t)p.W * (size_t)p.H;
    State s{n, x, v, C, F, J, mat};
    std::vector<double> grid(3 * G, 0.0);
    int nt = 1;
#ifdef _OPENMP
    if (threads > 1) { nt = threads; }
#else
    (void)threads;
#endif
    std::vector<double> priv;
    if (nt > 1) { priv.assign((size_t)nt * 3 * G, 0.0); }
    for (int step = 0; step < substeps; ++step) {
        std::fill(grid.begin(), grid.end(), 0.0);  // resetGridKernel, MPM.cu:46-51
        if (nt == 1) {
            p2g_range(p, s, 0, n, grid.data());
            grid_update_range(p, grid.data(), 0, G);
            g2p_range(p, s, 0, n, grid.data());
        } else {
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
            {
                const int t = omp_get_thread_num();
                double* mine = &priv[(size_t)t * 3 * G];
                std::fill(mine, mine + 3 * G, 0.0);
                const int lo = (int)((int64_t)n * t / nt), hi = (int)((int64_t)n * (t + 1) / nt);
                p2g_range(p, s, lo, hi, mine);
#pragma omp barrier
                // Fixed-order reduction: node range per thread, sum over threads 0..nt-1 in order.
                const size_t glo = G * (size_t)t / (size_t)nt, ghi = G * (size_t)(t + 1) / (size_t)nt;
                for (size_t k = 3 * glo; k < 3 * ghi; ++k) {
                    double acc = 0.0;
                    for (int u = 0; u < nt; ++u) { acc += priv[(size_t)u * 3 * G + k]; }
                    grid[k] = acc;
                }
                grid_update_range(p, grid.data(), glo, ghi);
#pragma omp barrier
                g2p_range(p, s, lo, hi, grid.data());
            }
#endif
        }
    }
    if (grid_out) {
        for (size_t k = 0; k < G; ++k) {
            grid_out[k] = grid[3 * k + 2];
            grid_out[G + k] = grid[3 * k + 0];
            grid_out[2 * G + k] = grid[3 * k + 1];
        }
    }
    return 0;
}

inline double silu(double z) { return z / (1.0 + std::exp(-z)); }
inline double dsilu(double z)
{
    const double sg = 1.0 / (1.0 + std::exp(-z));
    return sg * (1.0 + z * (1.0 - sg));
}

}  // namespace

extern "C" {

// Full substeps of one scene, state in FP64 (updated in place).  threads <= 1: serial loop.
// grid_out (optional): float64[3*G] = mass plane, v_x plane, v_y plane of the last substep's grid.
int32_t oracle_run(const MpmParams* params, int32_t n, double* x, double* v, double* C, double* F, double* J,
                   const uint8_t* mat, int32_t substeps, int32_t threads, double* grid_out)
{
    return run(params, n, x, v, C, F, J, mat, substeps, threads, grid_out);
}

// Individual phases for per-kernel parity.  grid is float64[3*G] interleaved (mom_x, mom_y, mass) / (v_x, v_y, mass).
int32_t oracle_p2g(const MpmParams* params, int32_t n, const double* x, const double* v, const double* C,
                   const double* F, const double* J, const uint8_t* mat, double* grid)
{
    const P64 p = promote(*params);
    State s{n, const_cast<double*>(x), const_cast<double*>(v), const_cast<double*>(C), const_cast<double*>(F),
            const_cast<double*>(J), mat};
    p2g_range(p, s, 0, n, grid);
    return 0;
}
int32_t oracle_grid_update(const MpmParams* params, double* grid)
{
    const P64 p = promote(*params);
    grid_update_range(p, grid, 0, (size_t)p.W * (size_t)p.H);
    return 0;
}
int32_t oracle_g2p(const MpmParams* params, int32_t n, double* x, double* v, double* C, double* F, double* J,
                   const uint8_t* mat, const double* grid)
{
    const P64 p = promote(*params);
    State s{n, x, v, C, F, J, mat};
    g2p_range(p, s, 0, n, grid);
    return 0;
}

// Quadratic B-spline weights for one coordinate: base node, fraction, w[3].
void oracle_weights(const MpmParams* params, double x, int32_t* base, double* f, double* w)
{
    const P64 p = promote(*params);
    int b;
    bspline(x, p, b, *f, w);
    *base = b;
}

// Per-particle Kirchhoff stress tau[4*n].
void oracle_tau(const MpmParams* params, int32_t n, const uint8_t* mat, const double* F, const double* J, double* tau)
{
    const P64 p = promote(*params);
    for (int i = 0; i < n; ++i) { particle_tau(p, mat[i], &F[4 * i], J[i], &tau[4 * i]); }
}
void oracle_polar(int32_t n, const double* F, double* R)
{
    for (int i = 0; i < n; ++i) { polar_rotation(&F[4 * i], &R[4 * i]); }
}
void oracle_svd(int32_t n, const double* F, double* U, double* sig, double* V)
{
    for (int i = 0; i < n; ++i) { svd2(&F[4 * i], &U[4 * i], &sig[2 * i], &V[4 * i]); }
}
void oracle_snow_project(const MpmParams* params, int32_t n, double* F, double* Jp)
{
    const P64 p = promote(*params);
    for (int i = 0; i < n; ++i) { snow_project(p, &F[4 * i], Jp[i]); }
}

// FP32 bin key of each particle, computed as the engine does: base = floorf(fmaf(x, inv_dx, halo) - 0.5f),
// clamped to [0, dim-3]; key = scene * (nbx*nby) + (base_y / block_y) * nbx + base_x / block_x,
// nbx = ceil(grid_w / block_x).  scene_of: int32[n] or NULL (all scene 0).
void oracle_bin_keys_f32(const MpmParams* params, int32_t n, const float* x, const int32_t* scene_of, int32_t block_x,
                         int32_t block_y, uint32_t* keys)
{
    const int nbx = (params->grid_w + block_x - 1) / block_x;
    const int nby = (params->grid_h + block_y - 1) / block_y;
    const float h = (float)params->halo;
    for (int i = 0; i < n; ++i) {
        int bx = (int)std::floor(std::fmaf(x[2 * i], params->inv_dx, h) - 0.5f);
        int by = (int)std::floor(std::fmaf(x[2 * i + 1], params->inv_dx, h) - 0.5f);
        bx = std::min(std::max(bx, 0), params->grid_w - 3);
        by = std::min(std::max(by, 0), params->grid_h - 3);
        const uint32_t sc = scene_of ? (uint32_t)scene_of[i] : 0u;
        keys[i] = sc * (uint32_t)(nbx * nby) + (uint32_t)((by / block_y) * nbx + bx / block_x);
    }
}
// Stable ascending order of keys: perm[j] = index of the j-th particle after sorting.
void oracle_stable_order(int32_t n, const uint32_t* keys, int32_t* perm)
{
    std::vector<int32_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) { return keys[a] < keys[b]; });
    std::memcpy(perm, idx.data(), sizeof(int32_t) * (size_t)n);
}

// Stencil base node of each particle within its bin (row-major, x fastest), from the same FP32 base node
// as oracle_bin_keys_f32.
void oracle_bin_nodes_f32(const MpmParams* params, int32_t n, const float* x, int32_t block_x, int32_t block_y,
                          uint32_t* nodes)
{
    const float h = (float)params->halo;
    for (int i = 0; i < n; ++i) {
        int bx = (int)std::floor(std::fmaf(x[2 * i], params->inv_dx, h) - 0.5f);
        int by = (int)std::floor(std::fmaf(x[2 * i + 1], params->inv_dx, h) - 0.5f);
        bx = std::min(std::max(bx, 0), params->grid_w - 3);
        by = std::min(std::max(by, 0), params->grid_h - 3);
        nodes[i] = (uint32_t)((by % block_y) * block_x + bx % block_x);
    }
}
// Checkerboard slot of each in-bin node (ly * block_x + lx, block_x even): the nodes with lx + ly even take
// slots 0 .. nb/2 - 1 in node order, those with lx + ly odd slots nb/2 .. nb - 1, nb = block_x * block_y.
// Counted directly: slot = parity * nb/2 + (nodes of the same parity before it) = parity * nb/2 +
// ly * block_x/2 + lx / 2, since each row holds block_x/2 nodes of each parity.
void oracle_checker_slots(int32_t n, const uint32_t* nodes, int32_t block_x, int32_t block_y, uint32_t* slots)
{
    const uint32_t half = (uint32_t)(block_x * block_y / 2);
    for (int i = 0; i < n; ++i) {
        const uint32_t lx = nodes[i] % (uint32_t)block_x, ly = nodes[i] / (uint32_t)block_x;
        slots[i] = ((lx + ly) & 1u) * half + ly * (uint32_t)(block_x / 2) + lx / 2;
    }
}
// Layered binned order: particles grouped by bin key (ascending); inside a bin, the particle of rank r
// among those of its node (in input order) goes to layer min(r, layers - 1); layers in order, nodes in
// ascending label order inside a layer, and the last layer in (node, rank) order.  The engine labels
// nodes by their checkerboard slots (oracle_checker_slots).  perm[j] = input index of position j.
void oracle_layered_order(int32_t n, const uint32_t* keys, const uint32_t* nodes, int32_t layers, int32_t* perm)
{
    std::vector<int32_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) {
        return keys[a] != keys[b] ? keys[a] < keys[b] : nodes[a] < nodes[b];
    });
    std::vector<int32_t> rank(n);
    for (int j = 0; j < n; ++j) {
        const int32_t i = idx[j];
        const bool same = j > 0 && keys[idx[j - 1]] == keys[i] && nodes[idx[j - 1]] == nodes[i];
        rank[i] = same ? rank[idx[j - 1]] + 1 : 0;
    }
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) {
        if (keys[a] != keys[b]) { return keys[a] < keys[b]; }
        const int32_t la = std::min(rank[a], layers - 1), lb = std::min(rank[b], layers - 1);
        if (la != lb) { return la < lb; }
        if (nodes[a] != nodes[b]) { return nodes[a] < nodes[b]; }
        return rank[a] < rank[b];
    });
    std::memcpy(perm, idx.data(), sizeof(int32_t) * (size_t)n);
}

// ---- MLP closure restatement (FP64) ----
// Layout: W1[64][4], b1[64], W2[64][64], b2[64], W3[3][64], b3[3]  (row = output unit).
// forward: h1 = silu(W1 x + b1), h2 = silu(W2 h1 + b2), y = W3 h2 + b3.
void oracle_mlp_forward(int32_t n, const double* in /* n x 4 */, const double* W1, const double* b1,
                        const double* W2, const double* b2, const double* W3, const double* b3, double* out /* n x 3 */)
{
    for (int i = 0; i < n; ++i) {
        double z1[64], h1[64], z2[64], h2[64];
        for (int j = 0; j < 64; ++j) {
            double a = b1[j];
            for (int k = 0; k < 4; ++k) { a += W1[j * 4 + k] * in[4 * i + k]; }
            z1[j] = a;
            h1[j] = silu(a);
        }
        for (int j = 0; j < 64; ++j) {
            double a = b2[j];
            for (int k = 0; k < 64; ++k) { a += W2[j * 64 + k] * h1[k]; }
            z2[j] = a;
            h2[j] = silu(a);
        }
        for (int j = 0; j < 3; ++j) {
            double a = b3[j];
            for (int k = 0; k < 64; ++k) { a += W3[j * 64 + k] * h2[k]; }
            out[3 * i + j] = a;
        }
        (void)z1;
        (void)z2;
    }
}
// Loss L = (1/n) sum_i sum_j wgt_j (y_ij - t_ij)^2 and its gradient w.r.t. every parameter (accumulated into g*).
double oracle_mlp_loss_grad(int32_t n, const double* in, const double* target, const double* wgt /* 3 */,
                            const double* W1, const double* b1, const double* W2, const double* b2, const double* W3,
                            const double* b3, double* gW1, double* gb1, double* gW2, double* gb2, double* gW3,
                            double* gb3)
{
    double loss = 0.0;
    for (int i = 0; i < n; ++i) {
        double z1[64], h1[64], z2[64], h2[64], y[3];
        for (int j = 0; j < 64; ++j) {
            double a = b1[j];
            for (int k = 0; k < 4; ++k) { a += W1[j * 4 + k] * in[4 * i + k]; }
            z1[j] = a;
            h1[j] = silu(a);
        }
        for (int j = 0; j < 64; ++j) {
            double a = b2[j];
            for (int k = 0; k < 64; ++k) { a += W2[j * 64 + k] * h1[k]; }
            z2[j] = a;
            h2[j] = silu(a);
        }
        double dy[3];
        for (int j = 0; j < 3; ++j) {
            double a = b3[j];
            for (int k = 0; k < 64; ++k) { a += W3[j * 64 + k] * h2[k]; }
            y[j] = a;
            const double e = y[j] - target[3 * i + j];
            loss += wgt[j] * e * e;
            dy[j] = 2.0 * wgt[j] * e / (double)n;
        }
        double dh2[64] = {0.0};
        for (int j = 0; j < 3; ++j) {
            gb3[j] += dy[j];
            for (int k = 0; k < 64; ++k) {
                gW3[j * 64 + k] += dy[j] * h2[k];
                dh2[k] += dy[j] * W3[j * 64 + k];
            }
        }
        double dh1[64] = {0.0};
        for (int j = 0; j < 64; ++j) {
            const double dz = dh2[j] * dsilu(z2[j]);
            gb2[j] += dz;
            for (int k = 0; k < 64; ++k) {
                gW2[j * 64 + k] += dz * h1[k];
                dh1[k] += dz * W2[j * 64 + k];
            }
        }
        for (int j = 0; j < 64; ++j) {
            const double dz = dh1[j] * dsilu(z1[j]);
            gb1[j] += dz;
            for (int k = 0; k < 4; ++k) { gW1[j * 4 + k] += dz * in[4 * i + k]; }
        }
    }
    return loss / (double)n;
}
// One Adam step (Kingma & Ba 2015) over a flat parameter vector; t is the 1-based step count.
void oracle_adam(int32_t n, double* w, const double* g, double* m, double* v, double lr, double beta1, double beta2,
                 double eps, int32_t t)
{
    const double c1 = 1.0 - std::pow(beta1, (double)t);
    const double c2 = 1.0 - std::pow(beta2, (double)t);
    for (int i = 0; i < n; ++i) {
        m[i] = beta1 * m[i] + (1.0 - beta1) * g[i];
        v[i] = beta2 * v[i] + (1.0 - beta2) * g[i] * g[i];
        const double mh = m[i] / c1, vh = v[i] / c2;
        w[i] -= lr * mh / (std::sqrt(vh) + eps);
    }
}

// Serial restatement of the reference's frame: backgroundKernel (libs/render/source/cuda-render.cu:214-243)
// over the whole image, then drawKernel (:259-310) for particles 0..n-1 in index order, each blending the
// colour (140, 210, 255, alpha 100) with alpha() (:244-257) over the 5x5 block minus corners around
// ((int)(x * scale), (int)(y * scale)).  FP32 arithmetic as written (this file is built without contraction).
// flip_extent > 0 draws y as flip_extent - y.  rgba: uint8[4 * width * height], rows packed.
void oracle_render(int32_t n, const float* pos, int32_t width, int32_t height, int32_t chunk_px, float scale,
                   float flip_extent, uint8_t* rgba)
{
    for (int y = 0; y < height; ++y) {
        for (int x = 0; x < width; ++x) {
            uint8_t* px = rgba + 4 * ((size_t)y * width + x);
            const bool line = chunk_px > 0 && (x % chunk_px == 0 || y % chunk_px == 0);
            if (line) {
                px[0] = 0;
                px[1] = 0;
                px[2] = 0;
            } else {
                const uint8_t shade = (uint8_t)(100.f * ((float)y / (float)height));
                px[0] = (uint8_t)(255u - shade);
                px[1] = (uint8_t)(160u - shade);
                px[2] = 80;
            }
            px[3] = 255;
        }
    }
    const float c = 1.f / 255.f;
    const uint8_t over[4] = {140, 210, 255, 100};
    const float a = (float)over[3] * c;
    const float ia = 1.f - a;
    for (int i = 0; i < n; ++i) {
        const float fx = pos[2 * i];
        float fy = pos[2 * i + 1];
        if (flip_extent > 0.f) { fy = flip_extent - fy; }
        const float sx = fx * scale, sy = fy * scale;
        if (!(sx > -2147483520.f && sx < 2147483520.f && sy > -2147483520.f && sy < 2147483520.f)) { continue; }
        const int32_t rx = (int32_t)sx, ry = (int32_t)sy;
        if ((size_t)(int64_t)rx >= (size_t)width || (size_t)(int64_t)ry >= (size_t)height) { continue; }
        for (int di = -2; di <= 2; ++di) {
            if (ry + di >= height || ry + di < 0) { continue; }
            for (int dj = -2; dj <= 2; ++dj) {
                if (rx + dj >= width || rx + dj < 0) { continue; }
                if (std::abs(di) == 2 && std::abs(dj) == 2) { continue; }
                uint8_t* px = rgba + 4 * ((size_t)(ry + di) * width + (rx + dj));
                for (int ch = 0; ch < 3; ++ch) {
                    px[ch] = (uint8_t)((float)over[ch] * a + (float)px[ch] * ia + 0.5f);
                }
                px[3] = 255;
            }
        }
    }
}

}  // extern "C"
