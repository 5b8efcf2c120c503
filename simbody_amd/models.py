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
TRANSLATION, PLANAR, CYLINDER, SCREW, GIMBAL, WELD, BENDSTRETCH, BUSHING = 5, 6, 7, 8, 9, 10, 11, 12
DOF = {PIN: 1, SLIDER: 1, UNIVERSAL: 2, BALL: 3, FREE: 6, TRANSLATION: 3, PLANAR: 3, CYLINDER: 2, SCREW: 1, GIMBAL: 3,
       WELD: 0, BENDSTRETCH: 2, BUSHING: 6}
NQ = {PIN: 1, SLIDER: 1, UNIVERSAL: 2, BALL: 4, FREE: 7, TRANSLATION: 3, PLANAR: 3, CYLINDER: 2, SCREW: 1,
      GIMBAL: 3, WELD: 0, BENDSTRETCH: 2, BUSHING: 6}
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
    pitch: np.ndarray = None  # [nb] Screw pitch (length per radian), None = zeros
    reversed: np.ndarray = None  # [nb] int32, nonzero = MobilizedBody::Reverse; None = all forward
    use_euler_angles: bool = False  # Ball/Free q as body-fixed x-y-z angles (4/7 q slots kept)

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


def stiff_chain(nb=32, seed=2, max_ratio=1e4, types=(PIN, UNIVERSAL, BALL)):
    """spatial_chain(nb, seed, types) with link masses (and, with them, the
    inertias about each body origin) scaled by max_ratio**u_b, u_b ~ U(0,1)
    per link, so neighbouring links differ in mass by ratios up to
    max_ratio."""
    spec = spatial_chain(nb, seed, types)
    f = max_ratio ** Rng(seed, 3).uniform(nb)
    spec.mass = spec.mass * f
    spec.inertia_B = spec.inertia_B * f[:, None]
    return spec


_TREE_TYPES = (PIN, SLIDER, UNIVERSAL, BALL)
_TREE_P = np.array([0.40, 0.10, 0.20, 0.25])


def _normals(U):
    """Box-Muller normals from pairs of uniform columns: U[..., 2k] and
    U[..., 2k+1] give normal k."""
    u1 = U[..., 0::2]; u2 = U[..., 1::2]
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def _quat_rows_to_rot(q):
    """Rotation matrices [n,3,3] of (not necessarily unit) quaternions [n,4]."""
    q = q / np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    return np.stack([np.stack([w*w + x*x - y*y - z*z, 2*(x*y - w*z), 2*(x*z + w*y)], -1),
                     np.stack([2*(x*y + w*z), w*w - x*x + y*y - z*z, 2*(y*z - w*x)], -1),
                     np.stack([2*(x*z - w*y), 2*(y*z + w*x), w*w - x*x - y*y + z*z], -1)], 1)


# Per-body counter layout of the bulk generators (uniforms per body):
# 0 parent, 1 mobilizer, 2-3 mass normal, 4 link length, 5-10 link direction
# (3 normals), 11-13 principal-moment draws, 14-21 principal axes (quaternion
# of 4 normals), 22-29 X_BM rotation, 30-37 X_PF rotation, 38-43 X_PF offset.
_BODY_DRAWS = 44


def _links_from_uniforms(U):
    """Mass properties and joint offsets of the links whose draws are the rows
    of U [n, _BODY_DRAWS]: mass ~ LogNormal(0, 0.5); central inertia with
    moments (y+z, x+z, x+y), x,y,z ~ U(0.02, 0.2) times m, on random principal
    axes; inboard joint at L*d in B with L ~ U(0.2, 1) and d a random unit
    vector; mass center at L*d/2.  Returns mass [n], com [n,3], inertia about
    Bo [n,6], p_BM [n,3]."""
    n = U.shape[0]
    m = np.exp(0.5 * _normals(U[:, 2:4])[:, 0])
    L = 0.2 + 0.8 * U[:, 4]
    d = _normals(U[:, 5:11]); d /= np.linalg.norm(d, axis=1, keepdims=True)
    p = L[:, None] * d
    c = 0.5 * p
    xyz = 0.02 + 0.18 * U[:, 11:14]
    mom = np.stack([xyz[:, 1] + xyz[:, 2], xyz[:, 0] + xyz[:, 2], xyz[:, 0] + xyz[:, 1]], 1)
    R = _quat_rows_to_rot(_normals(U[:, 14:22]))
    Ic = m[:, None, None] * np.einsum("nij,nj,nkj->nik", R, mom, R)
    cc = np.einsum("ni,ni->n", c, c)
    Io = Ic + m[:, None, None] * (cc[:, None, None] * np.eye(3)[None] - np.einsum("ni,nj->nij", c, c))
    I6 = np.stack([Io[:, 0, 0], Io[:, 1, 1], Io[:, 2, 2], Io[:, 0, 1], Io[:, 0, 2], Io[:, 1, 2]], 1)
    return m, c, I6, p


def random_tree(seed, index, nmin=8, nmax=128, max_depth=20, general_frames=False):
    """Uneven branched tree: nb ~ U{nmin..nmax}; body 1 is the root on Ground
    (depth 1); each later body's parent is uniform among the earlier bodies of
    depth < max_depth, so no body is deeper than max_depth.  The root draws
    Pin .40 / Slider .10 / Universal .20 / Ball .25 / Free .05; other bodies
    draw from the first four renormalized.  Mass, inertia and offsets as in
    spatial_chain.  With general_frames the X_PF and X_BM carry random
    rotations as well (parity tests).  All draws of a tree come from one
    counter stream in a fixed per-body layout."""
    rng = Rng(seed, 1000003 + index)
    head = rng.uniform(2)
    nb = min(nmax, nmin + int(head[0] * (nmax - nmin + 1)))
    U = rng.uniform(nb * _BODY_DRAWS).reshape(nb, _BODY_DRAWS)
    parent = np.zeros(nb, np.int32)
    depth = [0, 1]
    cand = [1] if max_depth > 1 else []
    for b in range(2, nb + 1):
        pb = cand[min(len(cand) - 1, int(U[b - 1, 0] * len(cand)))]
        parent[b - 1] = pb
        depth.append(depth[pb] + 1)
        if depth[b] < max_depth:
            cand.append(b)
    mob = np.zeros(nb, np.int32)
    cum = np.cumsum(np.append(_TREE_P, 0.05))
    mob[0] = [PIN, SLIDER, UNIVERSAL, BALL, FREE][int(np.searchsorted(cum, head[1] * cum[-1], side="right").clip(0, 4))]
    pn = np.cumsum(_TREE_P / _TREE_P.sum())
    if nb > 1:
        mob[1:] = np.array(_TREE_TYPES)[np.searchsorted(pn, U[1:, 1], side="right").clip(0, 3)]
    mass, com, inert, p = _links_from_uniforms(U)
    X_BM = np.zeros((nb, 12)); X_PF = np.zeros((nb, 12))
    X_BM[:, 9:] = p
    if general_frames:
        X_BM[:, :9] = _quat_rows_to_rot(_normals(U[:, 22:30])).reshape(nb, 9)
        X_PF[:, :9] = _quat_rows_to_rot(_normals(U[:, 30:38])).reshape(nb, 9)
        X_PF[:, 9:] = 0.1 * _normals(U[:, 38:44])
    else:
        X_BM[:, :9] = np.eye(3).reshape(9); X_PF[:, :9] = np.eye(3).reshape(9)
    return TreeSpec(parent=parent, mobilizer=mob, mass=mass, com_B=com, inertia_B=inert,
                    X_PF=X_PF, X_BM=X_BM)


def chain_params(seed, indices, nb=32, types=(PIN, UNIVERSAL, BALL)):
    """Per-instance parameters of chains with the topology of
    spatial_chain(nb, types=types) and the same distributions, one
    independent draw per instance index (the "each request carries its own
    tree" serving workload).  Returns mass [n,nb], com [n,nb,3], inertia
    [n,nb,6], X_PF [n,nb,12], X_BM [n,nb,12]; row i depends only on
    (seed, indices[i])."""
    idx = np.asarray(indices)
    U = _uniform_matrix(seed + 0x5EED, idx, nb * _BODY_DRAWS).reshape(len(idx) * nb, _BODY_DRAWS)
    mass, com, inert, p = _links_from_uniforms(U)
    n = len(idx)
    X_BM = np.zeros((n * nb, 12)); X_BM[:, :9] = np.eye(3).reshape(9); X_BM[:, 9:] = p
    X_PF = np.tile(_xform(), (n * nb, 1))
    return (mass.reshape(n, nb), com.reshape(n, nb, 3), inert.reshape(n, nb, 6),
            X_PF.reshape(n, nb, 12), X_BM.reshape(n, nb, 12))


def chain_spec(params, i, types=(PIN, UNIVERSAL, BALL)):
    """TreeSpec of instance i of chain_params output."""
    mass, com, inert, X_PF, X_BM = params
    nb = mass.shape[1]
    return TreeSpec(parent=np.arange(nb, dtype=np.int32),
                    mobilizer=np.array([types[k % len(types)] for k in range(nb)], np.int32),
                    mass=mass[i].copy(), com_B=com[i].copy(), inertia_B=inert[i].copy(),
                    X_PF=X_PF[i].copy(), X_BM=X_BM[i].copy())


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
        elif t == WELD:
            continue
        elif t in (TRANSLATION, PLANAR, CYLINDER, SCREW, GIMBAL, BENDSTRETCH, BUSHING):
            kinds = {TRANSLATION: ("trans",) * 3, PLANAR: ("angle", "trans", "trans"),
                     CYLINDER: ("angle", "trans"), SCREW: ("angle",),
                     GIMBAL: ("angle", "tilt", "angle"), BENDSTRETCH: ("angle", "trans"),
                     BUSHING: ("angle", "tilt", "angle", "trans", "trans", "trans")}[t]
            for k in kinds:
                plan.append((k, qi, c)); qi += 1; c += 1
        elif getattr(spec, "use_euler_angles", False):
            # x-y-z angles, (Free) translation, one unused q slot left at 0
            for k in ("angle", "tilt", "angle") + (("trans",) * 3 if t == FREE else ()):
                plan.append((k, qi, c)); qi += 1; c += 1
            qi += 1
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
    angles U(-qscale, qscale), slider/free translations and Gimbal middle
    angles U(-1,1), random unit quaternions, u ~ N(0, uscale^2).  Row i depends only on (seed, indices[i])."""
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
        elif kind == "tilt":     # Gimbal's middle angle, > 0.5 rad from gimbal lock
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
