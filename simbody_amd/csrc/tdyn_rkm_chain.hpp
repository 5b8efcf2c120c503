// Rollout kernels of the chain evaluator (tdyn_chain.hpp), instantiated in
// one translation unit per model-access form and precision
// (tdyn_chain_*.hip) so that they compile in parallel.
#pragma once
#include "tdyn_internal.hpp"
#include "tdyn_math.hpp"
#include "tdyn_chain.hpp"

namespace tdyn {

// adjustStepSize (AbstractIntegratorRep.cpp:448-502).  Returns success.
__device__ __forceinline__ bool adjust_step(double err, double acc, bool limited, double hmin, double hmax,
                                            double& hcur) {
    double hnew;
    if (!isfinite(err)) hnew = 0.1*hcur;
    else if (err == 0) hnew = 5.0*hcur;
    else hnew = 0.9*hcur*pow(acc/err, 0.25);
    if (hnew > hcur) { if (limited || hnew < 1.2*hcur) hnew = hcur; }
    if (hnew < hcur) { if (err <= acc) hnew = hcur; else hnew = fmin(hnew, 0.9*hcur); }
    hnew = fmin(hnew, 5.0*hcur);
    hnew = fmax(hnew, 0.1*hcur);
    if (hmin != -1) hnew = fmax(hnew, hmin);
    if (hmax != -1) hnew = fmin(hnew, hmax);
    const bool ok = hnew >= hcur;
    hcur = hnew;
    return ok;
}

// Rollouts of a uniform serial-chain batch with the Merson stages fused into
// the chain evaluator's sweeps (no separate vector passes in the common
// path).  The state ping-pongs between VY and VY1; a lane whose current state
// sits in VY1 when its launch ends copies it back to VY.
template <class PA, class K, class GS = K, class TS = GS>
__global__ void __launch_bounds__(64) k_rkm_chain(RkmArgs a) {
    using namespace chain;
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= a.n_inst) return;
    if (a.status[lane] != 0) return;
    const ChainEval<PA, K, GS, TS> ev(a, lane, 0);
    const int4 ta = ldi4(a.T.tree_a, 0);
    const int nb = ta.x, nq = ta.y, nu = ta.z, ny = nq + nu;
    const int4* body_a = a.T.body_a + ta.w;
    const int nquat = ldi4(a.T.tree_b, 0).y;
    const size_t NV = (size_t)a.ny_max * a.L;
    const int L = a.L;
    auto rows = [&](int v) { return Rows{a.vec + v*NV, L, (unsigned)lane}; };
    const Rows C0 = rows(VC0), C1 = rows(VC1);
    const Rows S0 = rows(VY), S1 = rows(VY1), F0 = rows(VF0), FA = rows(VFA), FB = rows(VFB), YS = rows(VYS),
               YT = rows(VYT), YE = rows(VYE);

    const RkmCfg& cfg = a.cfg;
    const double acc = cfg.accuracy, consTol = cfg.cons_tol;
    double hmin = cfg.min_h, hmax = cfg.max_h;
    if (cfg.fixed_h > 0) { hmin = hmax = cfg.fixed_h; }
    const double projLimit = fmax(2*consTol, sqrt(consTol));
    double t = a.t[lane], hcur = a.h[lane];
    const double tEnd = a.t_end[lane];
    int attempts = a.stats[4*lane], accepted = a.stats[4*lane+1], rejected = a.stats[4*lane+2],
        projections = a.stats[4*lane+3];
    int status = 0, steps = 0;
    bool flip = false;     // current state in S1
    NormSums ns0;

    while (steps < cfg.max_steps) {
        if (t >= tEnd) { status = 1; break; }
        // Every active lane has taken the same number of accepted steps in
        // this launch, so flip is wave-uniform; saying so keeps the row
        // bases in scalar registers.
        const bool fl = __builtin_amdgcn_readfirstlane(flip ? 1 : 0) != 0;
        const Rows Y = fl ? S1 : S0, Y1 = fl ? S0 : S1;
        RkRows r{Y, F0, FA, FB, YT, F0, Y1, YE, fl ? C1 : C0, fl ? C0 : C1};
        ev.template derivs_rk<0>(r, 0.0, ns0);
        bool ok = false;
        double t1 = t;
        do {
            bool limited = false;
            if (tEnd < t + 0.95*hcur) { limited = true; t1 = tEnd; }
            else if (tEnd > t + 1.001*hcur) t1 = t + hcur;
            else t1 = tEnd;
            if (!(t1 > t)) { status = -3; break; }
            if (++attempts > cfg.max_attempts) { status = -2; break; }
            const double h = t1 - t;
            r.st = YT; r.out = FA; ev.template derivs_rk<1>(r, h, ns0);
            ev.template derivs_rk<2>(r, h, ns0);
            r.out = FB; ev.template derivs_rk<3>(r, h, ns0);
            NormSums ns{0.0, 0.0, 0.0, 0.0};
            r.st = YS; r.out = FA; ev.template derivs_rk<4>(r, h, ns);
            const double qn = nq ? sqrt((ns.sq_other + ns.sq_quat) / nq) : 0.0;
            const double un = nu ? sqrt(ns.su / nu) : 0.0;
            const double e1 = qn >= un ? qn : un;
            double err = e1;
            bool converged = true;
            if (!(e1 > 16.0*acc) && nquat > 0) {
                const double qe = sqrt(ns.s2 / nquat);
                if (qe > projLimit) converged = false;
                else if (qe > consTol) {
                    // normalizeQuaternions (SimbodyMatterSubsystemRep.cpp:4448-4475) and
                    // the re-evaluated error norm of the projected estimate
                    ++projections;
                    double sqq = 0;
                    for (int b = 0; b < nb; ++b) {
                        const int4 ba = ldi4(body_a, b);
                        if (ba.y != BALL && ba.y != FREE) continue;
                        double qq[4], n2 = 0, e[4], d4[4];
                        for (int k = 0; k < 4; ++k) { qq[k] = Y1(ba.z + k); n2 += qq[k]*qq[k]; }
                        const double n = sqrt(n2);
                        for (int k = 0; k < 4; ++k) { qq[k] = qq[k] / n; Y1(ba.z + k) = qq[k]; }
                        double d = 0;
                        for (int k = 0; k < 4; ++k) d += YE(ba.z + k) * qq[k];
                        for (int k = 0; k < 4; ++k) { e[k] = YE(ba.z + k) - d*qq[k]; YE(ba.z + k) = e[k]; }
                        const d3 w = quat_dot_to_w(qq, e);
                        quat_dot(qq, w, d4);
                        sqq += d4[0]*d4[0] + d4[1]*d4[1] + d4[2]*d4[2] + d4[3]*d4[3];
                    }
                    const double qn2 = nq ? sqrt((ns.sq_other + sqq) / nq) : 0.0;
                    err = qn2 >= un ? qn2 : un;
                }
            }
            if (!converged) err = __builtin_inf();
            ok = adjust_step(err, acc, limited, hmin, hmax, hcur);
            if (!ok) ++rejected;
        } while (!ok);
        if (status < 0) break;
        t = t1;
        flip = !flip;
        ++accepted; ++steps;
    }
    if (flip) {
        for (int k = 0; k < ny; ++k) S0(k) = S1(k);
        if constexpr (!std::is_same<K, double>::value)
            for (int k = 0; k < ny; ++k) C0(k) = C1(k);
    }
    if (status == 0 && t >= tEnd) status = 1;
    a.t[lane] = t; a.h[lane] = hcur; a.status[lane] = status;
    a.stats[4*lane] = attempts; a.stats[4*lane+1] = accepted; a.stats[4*lane+2] = rejected; a.stats[4*lane+3] = projections;
}

// f = f(y) of the current rollout state (vector VY) into VF0, by the
// rollout's evaluator at its precision: the derivative the integrator uses.
template <class PA, class K, class GS = K, class TS = GS>
__global__ void __launch_bounds__(64) k_chain_f0(RkmArgs a) {
    using namespace chain;
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= a.n_inst) return;
    const ChainEval<PA, K, GS, TS> ev(a, lane, 0);
    const size_t NV = (size_t)a.ny_max * a.L;
    auto rows = [&](int v) { return Rows{a.vec + v*NV, a.L, (unsigned)lane}; };
    const RkRows r{rows(VY), rows(VF0), rows(VFA), rows(VFB), rows(VYT), rows(VF0), rows(VY1), rows(VYE),
                   rows(VC0), rows(VC1)};
    NormSums ns{0.0, 0.0, 0.0, 0.0};
    ev.template derivs_rk<0>(r, 0.0, ns);
}

// Launches k_rkm_chain<PA, K, GS, TS> (f0_only: k_chain_f0).
template <class PA, class K, class GS = K, class TS = GS>
hipError_t launch_chain_t(const RkmArgs& a, hipStream_t s, bool f0_only) {
    const dim3 grd((a.n_inst + 63)/64), blk(64);
    if (f0_only) k_chain_f0<PA, K, GS, TS><<<grd, blk, 0, s>>>(a);
    else k_rkm_chain<PA, K, GS, TS><<<grd, blk, 0, s>>>(a);
    return hipGetLastError();
}

} // namespace tdyn
