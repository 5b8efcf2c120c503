"""GPU parity of forward dynamics with known udot (prescribed motion at the
acceleration level and locks) with the oracle, FP64 at 1e-12 relative per
instance, through the C ABI: batches of different trees mixing every
mobilizer (padded generic records) and one shared tree (typed records),
each instance with its own flags and known udots."""
import numpy as np
import pytest

from common import ext_trees, rel_inf
from simbody_amd import models as M
from oracle_ctypes import OracleModel

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-12


def _umask(spec, presc):
    m = []
    for b, t in enumerate(spec.mobilizer, start=1):
        m += [bool(presc[b])] * M.DOF[int(t)]
    return np.array(m, bool)


@pytest.mark.parametrize("kind", ["ragged", "shared"])
@pytest.mark.parametrize("lock", [False, True])
def test_known_udot_matches_oracle(kind, lock):
    from simbody_amd.matter import MatterBatch
    if kind == "ragged":
        per = ext_trees(181, 12, nmin=1, nmax=24)
        B = MatterBatch(per, tree_of=np.arange(len(per), dtype=np.int32))
    else:
        spec = M.random_tree(181, 0, nmin=30, nmax=30, general_frames=True)
        spec.mobilizer[0] = M.FREE
        per = [spec] * 48
        B = MatterBatch([spec], n_instances=48)
    qs, us, fs, fm, masks, known = [], [], [], [], [], []
    for i, t in enumerate(per):
        q, u = M.random_state(t, 181, i); qs.append(q); us.append(u)
        fs.append(M.random_body_forces(t, 182, i))
        r = np.random.default_rng(18100 + i)
        fm.append(r.normal(size=t.nu))
        p = (r.uniform(size=t.nb + 1) < 0.4).astype(np.int32); p[0] = 0
        masks.append(p)
        known.append(np.zeros(t.nu) if lock else np.where(_umask(t, p), r.normal(size=t.nu), 0.0))
    ud, tau, A = B.calcAccelerationWithKnownUDot(B.pack(qs), B.pack(us), B.pack(fm), B.pack(fs),
                                                 B.pack_body_flags(masks), B.pack(known))
    uds, taus, As = B.split_u(ud), B.split_u(tau), B.split_b(A, 6)
    for i, t in enumerate(per):
        o_ud, o_tau, o_A, rc = OracleModel(t).forward_prescribed(qs[i], us[i], fm[i], fs[i], masks[i], known[i])
        assert rc == 0
        assert rel_inf(uds[i], o_ud) < TOL, i
        assert rel_inf(As[i], o_A) < TOL, i
        if np.abs(o_tau).max() > 0:
            assert rel_inf(taus[i], o_tau) < TOL, i
        else:
            assert np.all(np.asarray(taus[i]) == 0.0), i


def _known_case(kind):
    from simbody_amd.matter import MatterBatch
    if kind == "ragged":
        per = ext_trees(201, 10, nmin=3, nmax=24)
        B = MatterBatch(per, tree_of=np.arange(len(per), dtype=np.int32))
    else:
        spec = M.spatial_chain(16, seed=11)
        per = [spec] * 32
        B = MatterBatch([spec], n_instances=32)
    qs, us, fs, masks, known = [], [], [], [], []
    for i, t in enumerate(per):
        q, u = M.random_state(t, 201, i)
        r = np.random.default_rng(20100 + i)
        p = (r.uniform(size=t.nb + 1) < 0.3).astype(np.int32); p[0] = 0
        m = _umask(t, p)
        if i % 2 == 0:
            u = np.where(m, 0.0, u); k = np.zeros(t.nu)          # locks
        else:
            k = np.where(m, r.normal(size=t.nu), 0.0)            # constant known udot
        qs.append(q); us.append(u); masks.append(p); known.append(k)
        fs.append(M.random_body_forces(t, 202, i))
    return per, B, qs, us, fs, masks, known


@pytest.mark.parametrize("kind", ["ragged", "shared"])
def test_rollouts_with_locks_and_known_udot(kind):
    """Rollouts with known udot (locks on some instances' bodies, constant
    prescribed udot on others) on every evaluator that takes them: the tree
    sweeps and, for different trees, the level-parallel evaluator at fixed
    and automatic lanes.  The oracle's step counts and states within 1e-8;
    the chain evaluator and the one-lane streaming evaluator refuse such a
    batch."""
    from simbody_amd.matter import RungeKuttaMerson, TdynError
    per, B, qs, us, fs, masks, known = _known_case(kind)
    kn, ku = B.pack_body_flags(masks), B.pack(known)
    B.set_known_udot(kn, ku)
    rkm = RungeKuttaMerson(accuracy=1e-6)
    for mode, lanes in ((3, 1), (2, 0)):
        with pytest.raises(TdynError):       # (a batch of different trees refuses mode 2 at once)
            B.set_tree_lanes(lanes); B.set_evaluator(mode)
            B.stepTo(B.pack(qs), B.pack(us), rkm, 0.1, fapp=B.pack(fs))
    ref = [OracleModel(t).integrate_known(np.concatenate([qs[i], us[i]]), 0.1, 1e-6, masks[i], known[i],
                                          Fapp=fs[i]) for i, t in enumerate(per)]
    runs = ((1, 0, 1), (0, 0, 4), (3, 8, 4), (3, 0, 4)) if kind == "ragged" else ((1, 0, 1), (0, 0, 1), (3, 0, 1))
    for mode, lanes, active in runs:
        B.set_evaluator(mode); B.set_tree_lanes(lanes)
        q1, u1, st = B.stepTo(B.pack(qs), B.pack(us), rkm, 0.1, fapp=B.pack(fs))
        assert B.active_evaluator()[0] == active, (mode, lanes)
        q1s, u1s, st = B.split_q(q1), B.split_u(u1), st.cpu().numpy()
        for i, t in enumerate(per):
            y, s = ref[i]
            assert (s["accepted"], s["rejected"]) == tuple(st[i, 1:3]), (mode, lanes, i, s, st[i])
            assert rel_inf(np.concatenate([q1s[i], u1s[i]]), y) < 1e-8, (mode, lanes, i)
            lk = _umask(t, masks[i])
            if i % 2 == 0:
                assert np.array_equal(u1s[i][lk], us[i][lk]), (mode, lanes, i)     # locked u stay put


def test_autotune_with_known_udot():
    """The autotuner probes the level-parallel settings and the tree sweeps
    for a batch with known udot (not the one-lane streaming evaluator), and
    the rollout it sets up still matches the oracle."""
    from simbody_amd.matter import RungeKuttaMerson
    per, B, qs, us, fs, masks, known = _known_case("ragged")
    B.set_known_udot(B.pack_body_flags(masks), B.pack(known))
    rkm = RungeKuttaMerson(accuracy=1e-6)
    chosen = B.autotune(B.pack(qs), B.pack(us), rkm, 0.02, fapp=B.pack(fs))
    assert chosen[1] != 1 and chosen[4] == 0, chosen
    q1, u1, st = B.stepTo(B.pack(qs), B.pack(us), rkm, 0.1, fapp=B.pack(fs))
    assert B.active_evaluator()[0] in (1, 4)
    q1s, u1s, st = B.split_q(q1), B.split_u(u1), st.cpu().numpy()
    for i, t in enumerate(per):
        y, s = OracleModel(t).integrate_known(np.concatenate([qs[i], us[i]]), 0.1, 1e-6, masks[i], known[i],
                                              Fapp=fs[i])
        assert (s["accepted"], s["rejected"]) == tuple(st[i, 1:3]), (i, s, st[i])
        assert rel_inf(np.concatenate([q1s[i], u1s[i]]), y) < 1e-8, i


def test_autotune_cache_hit_outside_the_candidates_is_reprobed(tmp_path):
    """A cached choice that this batch cannot run (the one-lane streaming
    evaluator, stored for a batch of the same shape without known udot) is
    not applied: the autotuner probes again, and the rollout it sets up
    begins."""
    from simbody_amd.matter import RungeKuttaMerson
    per, B, qs, us, fs, masks, known = _known_case("ragged")
    B.set_known_udot(B.pack_body_flags(masks), B.pack(known))
    rkm = RungeKuttaMerson(accuracy=1e-6)
    cache = str(tmp_path / "tune.txt")
    B.autotune(B.pack(qs), B.pack(us), rkm, 0.02, fapp=B.pack(fs), cache=cache)
    key = open(cache).read().splitlines()[-1].split("\t")[0]
    assert "known1" in key
    with open(cache, "a") as f:
        f.write(key + "\t3 1 0 32\n")
    chosen = B.autotune(B.pack(qs), B.pack(us), rkm, 0.02, fapp=B.pack(fs), cache=cache)
    assert chosen[1] != 1 and chosen[4] == 0, chosen
    B.stepTo(B.pack(qs), B.pack(us), rkm, 0.05, fapp=B.pack(fs))
