"""Tree descriptions and seeded workload generators.

A tree is described by plain arrays (the `tdyn_tree_desc` of include/tdyn.h):
bodies are numbered 1..nb with Ground = 0 and parent[b-1] < b, the order in
which Simbody creates MobilizedBodies (SimbodyMatterSubsystemRep.cpp:256-293,
q and u slots in that order).  Mobilizer codes: PIN 0, SLIDER 1, UNIVERSAL 2,
BALL 3 (quaternion), FREE 4 (quaternion + translation).  Inertias are about
the body origin, expressed in the body frame (Simbody's MassProperties
convention), stored as xx yy zz xy xz yz.  Transforms are 12 doubles:
row-major rotation then translation.

The generators use a counter-based hash (splitmix64 of seed, stream and
counter) so that every input is a pure function of (seed, instance index):
the GPU path, the oracle and the CPU baseline all see identical inputs.
"""
from dataclasses import dataclass, field

import numpy as np

PIN, SLIDER, UNIVERSAL, BALL, FREE = 0, 1, 2, 3, 4
DOF = {PIN: 1, SLIDER: 1, UNIVERSAL: 2, BALL: 3, FREE: 6}
NQ = {PIN: 1, SLIDER: 1, UNIVERSAL: 2, BALL: 4, FREE: 7}
NAMES = {PIN: "Pin", SLIDER: "Slider", UNIVERSAL: "Universal", BALL: "Ball", FREE: "Free"}


@dataclass
class TreeSpec:
    parent: np.ndarray      # int32 [nb]
    mobilizer: np.ndarray   # int32 [nb]
    mass: np.ndarray        # [nb]
    com_B: np.ndarray       # [nb,3]
    inertia_B: np.ndarray   # [nb,6]
    X_PF: np.ndarray        # [nb,12]
    X_BM: np.ndarray        # [nb,12]
    gravity: np.ndarray = field(default_factory=lambda: np.array([0.0, -9.8, 0.0]))

    @property
    def nb(self):
        return len(self.parent)

    @property
    def nq(self):
        return int(sum(NQ[int(t)] for t in self.mobilizer))

    @property
    def nu(self):
        return int(sum(DOF[int(t)] for t in self.mobilizer))

    def dofs(self):
        return [DOF[int(t)] for t in self.mobilizer]

    def offsets(self):
        """(qIndex, uIndex) per body in body-number order."""
        qi, ui, q, u = [], [], 0, 0
        for t in self.mobilizer:
            qi.append(q); ui.append(u)
            q += NQ[int(t)]; u += DOF[int(t)]
        return np.array(qi), np.array(ui)

    def depth(self):
        d = np.zeros(self.nb + 1, dtype=np.int64)
        for b in range(1, self.nb + 1):
            d[b] = d[self.parent[b - 1]] + 1
        return d[1:]


# ----------------------------------------------------------------------------
# Counter-based random numbers.
# ----------------------------------------------------------------------------
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return x ^ (x >> np.uint64(31))


class Rng:
    """uniform/normal draws that depend only on (seed, stream, counter)."""

    def __init__(self, seed, stream=0):
        with np.errstate(over="ignore"):
            self.key = _splitmix(np.uint64(seed) * np.uint64(0x100000001B3) ^ _splitmix(np.uint64(stream)))
        self.ctr = 0

    def _bits(self, n):
        c = np.arange(self.ctr, self.ctr + n, dtype=np.uint64)
        self.ctr += n
        with np.errstate(over="ignore"):
            return _splitmix(c ^ self.key)

    def uniform(self, n, lo=0.0, hi=1.0):
        u = (self._bits(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n):
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)

    def integers(self, n, lo, hi):          # inclusive range
        return (lo + np.floor(self.uniform(n) * (hi - lo + 1))).astype(np.int64)


def _xform(R=None, p=(0.0, 0.0, 0.0)):
    out = np.zeros(12)
    out[:9] = (np.eye(3) if R is None else np.asarray(R)).reshape(9)
    out[9:] = p
    return out


def quat_to_rot(q):
    q = np.asarray(q, dtype=np.float64) / np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[w*w + x*x - y*y - z*z, 2*(x*y - w*z), 2*(x*z + w*y)],
                     [2*(x*y + w*z), w*w - x*x + y*y - z*z, 2*(y*z - w*x)],
                     [2*(x*z - w*y), 2*(y*z + w*x), w*w - x*x - y*y + z*z]])


def _random_rotation(rng):
    return quat_to_rot(rng.normal(4))


def _unit_vector(rng):
    v = rng.normal(3)
    return v / np.linalg.norm(v)


def _random_inertia(rng, m):
    """Central inertia with positive principal moments that satisfy the
    triangle inequality (A=y+z, B=x+z, C=x+y with x,y,z>0), random axes."""
    x, y, z = rng.uniform(3, 0.02, 0.2)
    moments = np.array([y + z, x + z, x + y])
    R = _random_rotation(rng)
    return m * (R @ np.diag(moments) @ R.T)


def _about_origin(m, c, Ic):
    """Parallel-axis shift of a central inertia to the body origin."""
    return Ic + m * (np.dot(c, c) * np.eye(3) - np.outer(c, c))


def _sym6(I):
    return np.array([I[0, 0], I[1, 1], I[2, 2], I[0, 1], I[0, 2], I[1, 2]])


def _link(rng, m):
    """One link: offset L*d with L~U(0.2,1.0), d a random unit vector.  The
    inboard joint sits at L*d in the body frame (X_BM translation), the mass
    center halfway between the body origin and the joint."""
    L = rng.uniform(1, 0.2, 1.0)[0]
    d = _unit_vector(rng)
    c = 0.5 * L * d
    Io = _about_origin(m, c, _random_inertia(rng, m))
    return L * d, c, Io


def pendulum_chain(n=10):
    """n-link planar Pin chain about Z: mass 1, UnitInertia(1), mass center at
    the body origin, X_PF = identity, X_BM = translation (0,1,0), gravity 9.8
    along -Y (the README double pendulum extended to n links)."""
    return TreeSpec(parent=np.arange(n, dtype=np.int32), mobilizer=np.full(n, PIN, np.int32),
                    mass=np.ones(n), com_B=np.zeros((n, 3)),
                    inertia_B=np.tile([1.0, 1.0, 1.0, 0.0, 0.0, 0.0], (n, 1)),
                    X_PF=np.tile(_xform(), (n, 1)), X_BM=np.tile(_xform(p=(0, 1, 0)), (n, 1)),
                    gravity=np.array([0.0, -9.8, 0.0]))


def spatial_chain(nb=32, seed=2, types=(PIN, UNIVERSAL, BALL)):
    """Serial chain with mobilizer types cycling `types`; link mass ~
    LogNormal(0, 0.5); random principal axes; offsets U(0.2,1.0) along random
    unit vectors.  X_PF = identity (joint at the parent's origin)."""
    rng = Rng(seed, 1)
    mass = np.exp(0.5 * rng.normal(nb))
    com = np.zeros((nb, 3)); inert = np.zeros((nb, 6)); X_BM = np.zeros((nb, 12))
    for i in range(nb):
        p, c, Io = _link(rng, mass[i])
        com[i] = c; inert[i] = _sym6(Io); X_BM[i] = _xform(p=p)
    return TreeSpec(parent=np.arange(nb, dtype=np.int32),
                    mobilizer=np.array([types[i % len(types)] for i in range(nb)], np.int32),
                    mass=mass, com_B=com, inertia_B=inert,
                    X_PF=np.tile(_xform(), (nb, 1)), X_BM=X_BM)


_TREE_TYPES = (PIN, SLIDER, UNIVERSAL, BALL)
_TREE_P = np.array([0.40, 0.10, 0.20, 0.25])


def random_tree(seed, index, nmin=8, nmax=128, max_depth=20, general_frames=False):
    """Uneven branched tree: nb ~ U{nmin..nmax}; body 1 is the root on Ground;
    each later body's parent is uniform among earlier bodies of depth <
    max_depth.  The root draws Pin .40 / Slider .10 / Universal .20 / Ball .25 /
    Free .05; other bodies draw from the first four renormalized.  Mass,
    inertia and offsets as in spatial_chain.  With general_frames the X_PF and
    X_BM carry random rotations as well (parity tests)."""
    rng = Rng(seed, 1000003 + index)
    nb = int(rng.integers(1, nmin, nmax)[0])
    parent = np.zeros(nb, np.int32); depth = np.zeros(nb + 1, np.int64)
    for b in range(2, nb + 1):
        cand = np.nonzero(depth[1:b] < max_depth)[0] + 1
        parent[b - 1] = cand[int(rng.uniform(1)[0] * len(cand)) % len(cand)]
        depth[b] = depth[parent[b - 1]] + 1
    depth[1] = 1
    mob = np.zeros(nb, np.int32)
    r = rng.uniform(1)[0]
    cum = np.cumsum(np.append(_TREE_P, 0.05))
    mob[0] = [PIN, SLIDER, UNIVERSAL, BALL, FREE][int(np.searchsorted(cum, r * cum[-1], side="right").clip(0, 4))]
    pn = np.cumsum(_TREE_P / _TREE_P.sum())
    for b in range(1, nb):
        mob[b] = _TREE_TYPES[int(np.searchsorted(pn, rng.uniform(1)[0], side="right").clip(0, 3))]
    mass = np.exp(0.5 * rng.normal(nb))
    com = np.zeros((nb, 3)); inert = np.zeros((nb, 6)); X_BM = np.zeros((nb, 12)); X_PF = np.zeros((nb, 12))
    for i in range(nb):
        p, c, Io = _link(rng, mass[i])
        com[i] = c; inert[i] = _sym6(Io)
        if general_frames:
            X_BM[i] = _xform(_random_rotation(rng), p)
            X_PF[i] = _xform(_random_rotation(rng), 0.1 * rng.normal(3))
        else:
            X_BM[i] = _xform(p=p); X_PF[i] = _xform()
    return TreeSpec(parent=parent, mobilizer=mob, mass=mass, com_B=com, inertia_B=inert,
                    X_PF=X_PF, X_BM=X_BM)


def _state_plan(spec):
    """Counter layout of the per-instance state draws (one uniform per
    counter).  Returns a list of (kind, q_or_u, out_index, counters, args)."""
    plan, c, qi, ui = [], 0, 0, 0
    for t in spec.mobilizer:
        t = int(t)
        if t in (PIN, UNIVERSAL):
            for _ in range(DOF[t]):
                plan.append(("angle", qi, c)); qi += 1; c += 1
        elif t == SLIDER:
            plan.append(("trans", qi, c)); qi += 1; c += 1
        else:
            plan.append(("quat", qi, c)); qi += 4; c += 8
            if t == FREE:
                for _ in range(3):
                    plan.append(("trans", qi, c)); qi += 1; c += 1
        plan.append(("speed", ui, c, DOF[t])); ui += DOF[t]; c += 2 * DOF[t]
    return plan, c


def _uniform_matrix(seed, indices, n):
    """U[i, k] = the k-th uniform of Rng(seed, 7000001 + indices[i])."""
    idx = np.asarray(indices, dtype=np.uint64)
    with np.errstate(over="ignore"):
        keys = _splitmix(np.uint64(seed) * np.uint64(0x100000001B3) ^ _splitmix(np.uint64(7000001) + idx))
        bits = _splitmix(np.arange(n, dtype=np.uint64)[None, :] ^ keys[:, None])
    return (bits >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def random_states(spec, seed, indices, qscale=np.pi, uscale=1.0):
    """Initial states of instances `indices`, instance-major (n, nq + nu):
    angles U(-qscale, qscale), slider/free translations U(-1,1), random unit
    quaternions, u ~ N(0, uscale^2).  Row i depends only on (seed, indices[i])."""
    plan, nc = _state_plan(spec)
    nq = spec.nq
    U = _uniform_matrix(seed, indices, nc)
    Y = np.zeros((U.shape[0], nq + spec.nu))

    def normal(c, k):       # k normals: first k draws are u1, next k are u2 (Rng.normal)
        u1 = U[:, c:c + k]; u2 = U[:, c + k:c + 2 * k]
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)

    for item in plan:
        kind, off, c = item[0], item[1], item[2]
        if kind == "angle":
            Y[:, off] = -qscale + 2 * qscale * U[:, c]
        elif kind == "trans":
            Y[:, off] = -1.0 + 2.0 * U[:, c]
        elif kind == "quat":
            qq = normal(c, 4)
            Y[:, off:off + 4] = qq / np.linalg.norm(qq, axis=1, keepdims=True)
        else:
            d = item[3]
            Y[:, nq + off:nq + off + d] = uscale * normal(c, d)
    return Y


def random_state(spec, seed, index, qscale=np.pi, uscale=1.0):
    """(q, u) of one instance; see random_states."""
    y = random_states(spec, seed, [index], qscale, uscale)[0]
    return y[:spec.nq].copy(), y[spec.nq:].copy()


def random_body_forces(spec, seed, index):
    """Constant applied body forces in Ground, (nb+1) x 6 = (torque, force),
    each component ~ N(0,1); the Ground row is zero."""
    rng = Rng(seed, 9000011 + index)
    F = np.zeros((spec.nb + 1, 6))
    F[1:] = rng.normal(6 * spec.nb).reshape(spec.nb, 6)
    return F


def chain_batch_states(spec, n_inst, seed, chunk=8192):
    """Initial states for n_inst instances of one model, instance-major
    (n_inst, nq + nu)."""
    out = np.zeros((n_inst, spec.nq + spec.nu))
    for s in range(0, n_inst, chunk):
        e = min(n_inst, s + chunk)
        out[s:e] = random_states(spec, seed, np.arange(s, e))
    return out
