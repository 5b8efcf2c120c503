// Example caller of the C++ drop-in (include/tdyn_matter.hpp): the README
// pendulum of the reference extended to 10 Pin links, written the
// way a Simbody user writes it -- build the tree, realize(Acceleration), then
// TimeStepper::stepTo with RungeKuttaMerson at accuracy 1e-8.  Also prints
// M*udot (multiplyByM) and the composite mass of body 1
// (calcCompositeBodyInertias) at the initial state, the events of a
// reporter on the tip's x coordinate over the same run, and whether that
// run resumed from a checkpoint ends in the same state.
// Usage: dropin_example q0..q9 u0..u9  -> prints one JSON line.
#include "../../include/tdyn_matter.hpp"

#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <vector>

int main(int argc, char** argv) {
    const int n = 10;
    if (argc != 1 + 2 * n) { std::fprintf(stderr, "usage: %s q[10] u[10]\n", argv[0]); return 2; }
    std::vector<int32_t> parent(n), mob(n, TDYN_PIN);
    std::vector<double> mass(n, 1.0), com(3 * n, 0.0), inertia(6 * n, 0.0), X_PF(12 * n, 0.0), X_BM(12 * n, 0.0);
    for (int i = 0; i < n; ++i) {
        parent[i] = i;                                           // chain: body i+1 hangs from body i
        inertia[6*i] = inertia[6*i+1] = inertia[6*i+2] = 1.0;    // UnitInertia(1), mass 1
        X_PF[12*i] = X_PF[12*i+4] = X_PF[12*i+8] = 1.0;          // joint at the parent's origin
        X_BM[12*i] = X_BM[12*i+4] = X_BM[12*i+8] = 1.0; X_BM[12*i+10] = 1.0;   // (0,1,0) in B
    }
    tdyn_tree_desc d{n, parent.data(), mob.data(), mass.data(), com.data(), inertia.data(), X_PF.data(), X_BM.data(),
                     {0.0, -9.8, 0.0}};
    std::vector<double> hq(n), hu(n);
    for (int i = 0; i < n; ++i) { hq[i] = std::atof(argv[1 + i]); hu[i] = std::atof(argv[1 + n + i]); }
    double *q, *u, *udot, *Mud, *R; int32_t* stats;
    if (hipMalloc((void**)&q, n * 8) || hipMalloc((void**)&u, n * 8) || hipMalloc((void**)&udot, n * 8) ||
        hipMalloc((void**)&Mud, n * 8) || hipMalloc((void**)&R, 10 * (n + 1) * 8) ||
        hipMalloc((void**)&stats, 16)) { std::fprintf(stderr, "hipMalloc failed\n"); return 1; }
    (void)hipMemcpy(q, hq.data(), n * 8, hipMemcpyHostToDevice);
    (void)hipMemcpy(u, hu.data(), n * 8, hipMemcpyHostToDevice);
    try {
        tdyn::MatterBatch matter(d, 1);
        matter.realizeAcceleration(q, u, nullptr, udot);
        std::vector<double> hud(n), hMud(n), hR(10 * (n + 1));
        (void)hipMemcpy(hud.data(), udot, n * 8, hipMemcpyDeviceToHost);
        matter.multiplyByM(q, udot, Mud);
        matter.calcCompositeBodyInertias(q, R);
        (void)hipMemcpy(hMud.data(), Mud, n * 8, hipMemcpyDeviceToHost);
        (void)hipMemcpy(hR.data(), R, 10 * (n + 1) * 8, hipMemcpyDeviceToHost);
        tdyn::RungeKuttaMerson integ; integ.setAccuracy(1e-8);
        matter.stepTo(q, u, nullptr, 0.0, 1.0, integ, stats);
        std::vector<int32_t> hs(4);
        (void)hipMemcpy(hq.data(), q, n * 8, hipMemcpyDeviceToHost);
        (void)hipMemcpy(hu.data(), u, n * 8, hipMemcpyDeviceToHost);
        (void)hipMemcpy(hs.data(), stats, 16, hipMemcpyDeviceToHost);
        auto arr = [](const std::vector<double>& v) {
            std::string s = "[";
            for (size_t i = 0; i < v.size(); ++i) { char b[40]; std::snprintf(b, sizeof b, "%s%.17g", i ? "," : "", v[i]); s += b; }
            return s + "]";
        };
        // a TriggeredEventReporter on the x coordinate of body 10's origin,
        // and a checkpoint after 20 steps resumed in a second batch
        tdyn_witness_desc w{};
        w.kind = TDYN_WIT_STATION_POS; w.body = n; w.direction[0] = 1.0;
        w.trigger = TDYN_TRIGGER_FALLING | TDYN_TRIGGER_RISING; w.action = TDYN_EVENT_REPORT; w.window = 0.1;
        matter.setWitnesses(0, 1, &w);
        std::vector<double> q0(n), u0(n);
        for (int i = 0; i < n; ++i) { q0[i] = std::atof(argv[1 + i]); u0[i] = std::atof(argv[1 + n + i]); }
        (void)hipMemcpy(q, q0.data(), n * 8, hipMemcpyHostToDevice);
        (void)hipMemcpy(u, u0.data(), n * 8, hipMemcpyHostToDevice);
        matter.initialize(q, u, nullptr, 0.0, nullptr, 1.0, integ);
        matter.advance(20);
        const std::vector<unsigned char> snap = matter.snapshot();
        tdyn::MatterBatch resumed(d, 1);
        resumed.setWitnesses(0, 1, &w);
        resumed.restore(snap);
        matter.advance(1 << 20);
        resumed.advance(1 << 20);
        double* ev; int32_t *ids, *cnt;
        if (hipMalloc((void**)&ev, 64 * 2 * 8) || hipMalloc((void**)&ids, 64 * 2 * 4) || hipMalloc((void**)&cnt, 4)) {
            std::fprintf(stderr, "hipMalloc failed\n"); return 1;
        }
        matter.getEvents(ev, ids, cnt);
        std::vector<double> he(128); std::vector<int32_t> hi(128); int32_t hc = 0;
        (void)hipMemcpy(he.data(), ev, 128 * 8, hipMemcpyDeviceToHost);
        (void)hipMemcpy(hi.data(), ids, 128 * 4, hipMemcpyDeviceToHost);
        (void)hipMemcpy(&hc, cnt, 4, hipMemcpyDeviceToHost);
        std::vector<double> qa(n), qb(n), ua(n), ub(n);
        matter.getState(q, u, nullptr, nullptr, nullptr);
        (void)hipMemcpy(qa.data(), q, n * 8, hipMemcpyDeviceToHost); (void)hipMemcpy(ua.data(), u, n * 8, hipMemcpyDeviceToHost);
        resumed.getState(q, u, nullptr, nullptr, nullptr);
        (void)hipMemcpy(qb.data(), q, n * 8, hipMemcpyDeviceToHost); (void)hipMemcpy(ub.data(), u, n * 8, hipMemcpyDeviceToHost);
        const bool same = qa == qb && ua == ub;
        std::vector<double> win(he.begin(), he.begin() + 2 * std::min(hc, 64));
        std::vector<double> tr;
        for (int k = 0; k < std::min(hc, 64); ++k) tr.push_back(hi[2*k + 1]);
        std::printf("{\"udot\": %s, \"Mudot\": %s, \"mass1\": %.17g, \"q1\": %s, \"u1\": %s, "
                    "\"attempts\": %d, \"accepted\": %d, \"rejected\": %d, \"events\": %d, \"windows\": %s, "
                    "\"transitions\": %s, \"q_events\": %s, \"resume_same\": %s}\n",
                    arr(hud).c_str(), arr(hMud).c_str(), hR[10], arr(hq).c_str(), arr(hu).c_str(), hs[0], hs[1], hs[2],
                    hc, arr(win).c_str(), arr(tr).c_str(), arr(qa).c_str(), same ? "true" : "false");
    } catch (const tdyn::Error& e) {
        std::fprintf(stderr, "tdyn error %d: %s\n", e.code(), e.what());
        return 1;
    }
    return 0;
}
