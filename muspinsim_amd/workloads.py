"""Seeded synthetic workloads (orientation x field sweeps) used by the
benchmark and the parity tests.

Each builder returns (systems, InstanceBatch) in the reference's config order
(muspinsim/simconfig.py:300-316: file ranges x averaged axes x x-axis), with
orientation weights normalised as simconfig.py:152-159 does.

  c1_zero_field_powder   mu, e, 2 x 1H (n=16); B=0, 100 orientations, 1000 t in 0-10 us
  c2_alc_sweep           mu, e, 4 x 1H, 14N (n=192); 200 fields 0-3 T x 500 orientations,
                         integral (ALC preset)
  c3_tf_stress           mu + 8 19F/1H dipolar (n=512); 0.01 T, p = x, 256 orientations,
                         4000 t in 0-20 us
  c4_uneven_batch        4096 systems, n = 4 .. 1024, 101 t in 0-10 us
  c4_dissipative         1024 systems mu, e, 1H (n=8, Liouville N=64) with electron
                         dissipation a ~ U(1, 100) MHz, 2000 t in 0-10 us
  c5_fit                 mu, e, 4 x 1H (n=64); 256 orientations, 2000 t in 0-10 us
"""

import numpy as np

from .constants import ELEC_GAMMA, MU_GAMMA, dipolar_tensor, gyromagnetic_ratio
from .engine import InstanceBatch, SystemDesc

G_H = gyromagnetic_ratio("H")
G_F = gyromagnetic_ratio("F")
G_N = gyromagnetic_ratio("N")


def spiral_orientations(N):
    """N near-uniform polar angle pairs (theta, phi) on the sphere (a
    Zaremba-Conroy-Wolfsberg-style generalised spiral with equal weights)."""
    j = np.arange(N)
    theta = np.arccos(1.0 - 2.0 * (j + 0.5) / N)
    phi = np.mod(j * np.pi * (3.0 - np.sqrt(5.0)), 2 * np.pi)
    return np.stack([theta, phi], axis=1)


def polar_quats_conj(angles):
    """Stored (conjugated) quaternions of polar orientations (theta, phi):
    conj(Qz(phi) Qy(theta) Qz(phi)) (muspinsim/utils.py:53-68, simconfig.py:596-613)."""
    th, ph = angles[:, 0], angles[:, 1]
    q = np.empty((len(th), 4))
    for k in range(len(th)):
        qa = np.array([np.cos(ph[k] / 2), 0, 0, np.sin(ph[k] / 2)])
        qb = np.array([np.cos(th[k] / 2), 0, np.sin(th[k] / 2), 0])
        q[k] = _qmul(_qmul(qa, qb), qa)
    q[:, 1:] *= -1.0
    return q


def _qmul(p, q):
    w1, x1, y1, z1 = p
    w2, x2, y2, z2 = q
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
                     w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                     w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def _sym_aniso(rng, sigma):
    G = rng.normal(0.0, sigma, size=(3, 3))
    G = 0.5 * (G + G.T)
    return G - np.trace(G) / 3.0 * np.eye(3)


def _random_vector(rng, rmin, rmax):
    v = rng.normal(size=3)
    v /= np.linalg.norm(v)
    return v * rng.uniform(rmin, rmax)


def _rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _batch(q, B, p, w, slot, sys_id, times, y, n_slots, avg_N, T=np.inf, Bint=None):
    n = len(w)
    return InstanceBatch(
        q=np.ascontiguousarray(q, dtype=float), B=np.ascontiguousarray(B, dtype=float),
        Bint=np.zeros((n, 3)) if Bint is None else np.ascontiguousarray(Bint, dtype=float),
        p=np.ascontiguousarray(p, dtype=float), T=np.full(n, T, dtype=float),
        w=np.ascontiguousarray(w, dtype=float), slot=np.ascontiguousarray(slot, dtype=np.int64),
        sys_id=np.ascontiguousarray(sys_id, dtype=np.int64),
        times=np.ascontiguousarray(times, dtype=float), y=y, n_slots=n_slots, avg_N=avg_N)


def powder_batch(nori, fields, times, y, p, sys_ids=None):
    """Config product (orientation averaged, x = field or time): instances in
    orientation-major order, slots = field index (simconfig.py:300-316)."""
    ang = spiral_orientations(nori)
    q = polar_quats_conj(ang)
    nf = len(fields)
    qq = np.repeat(q, nf, axis=0)
    BB = np.tile(np.asarray(fields, dtype=float).reshape(nf, 3), (nori, 1))
    pp = np.tile(np.asarray(p, dtype=float), (nori * nf, 1))
    w = np.ones(nori * nf)                      # w * (N / sum w) with equal weights
    if y == "integral":
        slot = np.tile(np.arange(nf), nori)
        n_slots = nf
    else:
        slot = np.zeros(nori * nf, dtype=np.int64)
        n_slots = 1
    sid = np.zeros(nori * nf, dtype=np.int64) if sys_ids is None else sys_ids
    return _batch(qq, BB, pp, w, slot, sid, times, y, n_slots, nori)


def c1_zero_field_powder(seed=1, nori=100, ntimes=1000):
    rng = np.random.default_rng(seed)
    Is = [0.5, 0.5, 0.5, 0.5]
    g = [MU_GAMMA, ELEC_GAMMA, G_H, G_H]
    A = 100.0 * np.eye(3) + _sym_aniso(rng, 10.0)
    terms = [("bilinear", 0, 1, A)]
    for k in (2, 3):
        r = _random_vector(rng, 1.0, 3.0)
        terms.append(("bilinear", 0, k, dipolar_tensor(g[0], g[k], r)))
        terms.append(("bilinear", 1, k, dipolar_tensor(g[1], g[k], r)))
    sysd = SystemDesc(Is, g, terms, 0)
    times = np.linspace(0.0, 10.0, ntimes)
    return [sysd], powder_batch(nori, [[0, 0, 0]], times, "asymmetry", [0, 0, 1.0])


def c2_system(seed=2):
    rng = np.random.default_rng(seed)
    Is = [0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 1.0]
    g = [MU_GAMMA, ELEC_GAMMA, G_H, G_H, G_H, G_H, G_N]
    A = rng.uniform(200.0, 600.0) * np.eye(3) + _sym_aniso(rng, 20.0)
    terms = [("bilinear", 0, 1, A)]
    # nuclear hyperfine couplings to the electron (builder's choice, recorded)
    for k in range(2, 6):
        terms.append(("bilinear", k, 1, rng.uniform(1.0, 20.0) * np.eye(3) + _sym_aniso(rng, 2.0)))
    terms.append(("bilinear", 6, 1, rng.uniform(1.0, 10.0) * np.eye(3) + _sym_aniso(rng, 1.0)))
    # 14N quadrupole from Cq, eta in a random principal frame: I.Q.I with
    # Q = Cq/(2I(2I-1)) diag((eta-1)/2, (-eta-1)/2, 1)
    Cq = rng.uniform(1.0, 4.0)
    eta = rng.uniform(0.0, 1.0)
    I = 1.0
    Qp = Cq / (2 * I * (2 * I - 1)) * np.diag([(eta - 1) / 2, (-eta - 1) / 2, 1.0])
    R = _rotation(rng)
    terms.append(("bilinear", 6, 6, R @ Qp @ R.T))
    return SystemDesc(Is, g, terms, 0)


def c2_alc_sweep(seed=2, nfields=200, nori=500):
    sysd = c2_system(seed)
    fields = np.zeros((nfields, 3))
    fields[:, 2] = np.linspace(0.0, 3.0, nfields)
    return [sysd], powder_batch(nori, fields, np.zeros(1), "integral", [0, 0, 1.0])


def c3_tf_stress(seed=3, nori=256, ntimes=4000, nnuc=8):
    rng = np.random.default_rng(seed)
    kinds = ["F" if rng.uniform() < 0.5 else "H" for _ in range(nnuc)]
    Is = [0.5] * (nnuc + 1)
    g = [MU_GAMMA] + [G_F if k == "F" else G_H for k in kinds]
    pos = [np.zeros(3)] + [_random_vector(rng, 1.0, 2.5) for _ in range(nnuc)]
    terms = []
    for i in range(nnuc + 1):
        for j in range(i + 1, nnuc + 1):
            terms.append(("bilinear", i, j, dipolar_tensor(g[i], g[j], pos[j] - pos[i])))
    sysd = SystemDesc(Is, g, terms, 0)
    times = np.linspace(0.0, 20.0, ntimes)
    return [sysd], powder_batch(nori, [[0, 0, 0.01]], times, "asymmetry", [1.0, 0, 0])


def c4_uneven_batch(seed=4, ninst=4096, ntimes=101, max_dim=1024):
    rng = np.random.default_rng(seed)
    systems = []
    for _ in range(ninst):
        k = min(int(rng.geometric(0.35)) - 1, 8)
        spin1 = rng.uniform() < 0.3
        n = 4 * (2 ** k) * (3 if spin1 else 1)
        if n > max_dim:
            spin1 = False
        Is = [0.5, 0.5] + [0.5] * k + ([1.0] if spin1 else [])
        g = [MU_GAMMA, ELEC_GAMMA] + [G_H] * k + ([G_N] if spin1 else [])
        terms = [("bilinear", 0, 1, rng.uniform(50, 500) * np.eye(3) + _sym_aniso(rng, 10.0))]
        for i in range(2, len(Is)):
            terms.append(("bilinear", i, 1, rng.uniform(1, 20) * np.eye(3) + _sym_aniso(rng, 2.0)))
            r = _random_vector(rng, 1.0, 3.0)
            terms.append(("bilinear", 0, i, dipolar_tensor(g[0], g[i], r)))
        if spin1:
            Q = _sym_aniso(rng, 1.0)
            terms.append(("bilinear", len(Is) - 1, len(Is) - 1, Q))
        systems.append(SystemDesc(Is, g, terms, 0))
    ang = spiral_orientations(ninst)
    q = polar_quats_conj(ang)
    times = np.linspace(0.0, 10.0, ntimes)
    # one averaged output: every instance is its own orientation sample
    batch = _batch(q, np.zeros((ninst, 3)), np.tile([0, 0, 1.0], (ninst, 1)), np.ones(ninst),
                   np.zeros(ninst, dtype=np.int64), np.arange(ninst), times, "asymmetry", 1, ninst)
    return systems, batch


def c4_dissipative(seed=44, ninst=1024, ntimes=2000, B=(0.0, 0.0, 0.0), T=np.inf):
    """Dissipative subset: per-instance random hyperfine / dipolar couplings
    and electron dissipation rate, orientation averaged into one output."""
    rng = np.random.default_rng(seed)
    Is = [0.5, 0.5, 0.5]
    g = [MU_GAMMA, ELEC_GAMMA, G_H]
    systems = []
    for _ in range(ninst):
        terms = [("bilinear", 0, 1, rng.uniform(50, 500) * np.eye(3) + _sym_aniso(rng, 10.0)),
                 ("bilinear", 2, 1, rng.uniform(1, 20) * np.eye(3) + _sym_aniso(rng, 2.0))]
        r = _random_vector(rng, 1.0, 3.0)
        terms.append(("bilinear", 0, 2, dipolar_tensor(g[0], g[2], r)))
        systems.append(SystemDesc(Is, g, terms, 0, dissip={1: float(rng.uniform(1.0, 100.0))}))
    ang = spiral_orientations(ninst)
    q = polar_quats_conj(ang)
    times = np.linspace(0.0, 10.0, ntimes)
    batch = _batch(q, np.tile(np.asarray(B, dtype=float), (ninst, 1)),
                   np.tile([0, 0, 1.0], (ninst, 1)), np.ones(ninst),
                   np.zeros(ninst, dtype=np.int64), np.arange(ninst), times, "asymmetry", 1, ninst,
                   T=T)
    return systems, batch


def c5_system(params=None, seed=5):
    """mu, e, 4 x 1H; params = 6 hyperfine components (xx, yy, zz, xy, xz, yz)
    and 4 muon-proton distances (Angstrom)."""
    params, dirs = _c5_defaults(params, seed)
    a = params[:6]
    A = np.array([[a[0], a[3], a[4]], [a[3], a[1], a[5]], [a[4], a[5], a[2]]])
    Is = [0.5] * 6
    g = [MU_GAMMA, ELEC_GAMMA] + [G_H] * 4
    terms = [("bilinear", 0, 1, A)]
    for k in range(4):
        r = dirs[k] * params[6 + k]
        terms.append(("bilinear", 0, 2 + k, dipolar_tensor(g[0], g[2 + k], r)))
        # fixed isotropic proton-electron hyperfine (not a fit parameter)
        terms.append(("bilinear", 2 + k, 1, 10.0 * np.eye(3)))
    return SystemDesc(Is, g, terms, 0), params


def _c5_defaults(params, seed):
    """Default parameters and the (parameter-independent) proton directions."""
    rng = np.random.default_rng(seed)
    default = np.concatenate([[300.0, 320.0, 280.0, 10.0, -5.0, 8.0], rng.uniform(1.2, 2.5, 4)])
    dirs = rng.normal(size=(4, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    return (default if params is None else np.asarray(params, dtype=float)), dirs


C5_PARAMS = ("A_xx", "A_yy", "A_zz", "A_xy", "A_xz", "A_yz", "r_1", "r_2", "r_3", "r_4")


def c5_dterms(params, seed=5):
    """Derivative term lists dT/dtheta_p of the c5 system, one per parameter:
    unit symmetric tensors for the hyperfine components and -3/r_k times the
    muon-proton dipolar tensor for each distance (D ~ r^-3 at a fixed
    direction)."""
    params, dirs = _c5_defaults(params, seed)
    out = []
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    for a, b in pairs:
        E = np.zeros((3, 3))
        E[a, b] = 1.0
        E[b, a] = 1.0
        out.append([("bilinear", 0, 1, E)])
    for k in range(4):
        r = params[6 + k]
        D = dipolar_tensor(MU_GAMMA, G_H, dirs[k] * r)
        out.append([("bilinear", 0, 2 + k, -3.0 / r * D)])
    return out


def c5_fit(seed=5, nori=256, ntimes=2000):
    sysd, _ = c5_system(seed=seed)
    times = np.linspace(0.0, 10.0, ntimes)
    return [sysd], powder_batch(nori, [[0, 0, 0]], times, "asymmetry", [0, 0, 1.0])


def celio_mu4v_config(nori=256, averages=4, ntimes=50, k=10, field=0.001):
    """Celio workload: a muon among four 51V (I = 7/2, n = 8192) with
    muon-V dipolar couplings and V quadrupole tensors, a small longitudinal
    field so that the crystallite orientation matters, ZCW powder of at
    least `nori` orientations, `averages` random-phase states per
    orientation, times 0..10 us; the shape of the reference's celio example
    (examples/celio/celio.in: mu + 4 V, celio 10 4, 50 times) on a powder."""
    from .config import MuSpinConfig
    from .orientations import zcw
    from .spinsys import MuonSpinSystem
    s = MuonSpinSystem(["mu", "V", "V", "V", "V"], celio_k=k)
    r = [[0.0, -1.3, 1.2], [-1.25, 0.0, -1.22], [1.24, 0.0, -1.21], [0.0, 1.26, 1.23]]
    efg = [[[-0.031, 0, 0], [0, -0.009, -0.012], [0, -0.012, 0.040]],
           [[-0.009, 0, 0.012], [0, -0.030, 0], [0.012, 0, 0.039]],
           [[-0.009, 0, -0.012], [0, -0.031, 0], [-0.012, 0, 0.040]],
           [[-0.031, 0, 0], [0, -0.009, 0.012], [0, 0.012, 0.040]]]
    for j in range(4):
        s.add_bilinear_term(0, j + 1, _dip(r[j]), "Dipolar")
        s.add_bilinear_term(j + 1, j + 1, np.array(efg[j]), "Quadrupolar")
    return MuSpinConfig(s, field=field, orientation=zcw(nori), polarization="longitudinal",
                        time=np.linspace(0.0, 10.0, ntimes), celio=(k, averages))


def _dip(v):
    """A dipolar-like traceless tensor from a vector (MHz scale ~ |v|)."""
    v = np.asarray(v, dtype=float)
    u = v / np.linalg.norm(v)
    return np.linalg.norm(v) * (np.eye(3) - 3.0 * np.outer(u, u)) / 3.0
