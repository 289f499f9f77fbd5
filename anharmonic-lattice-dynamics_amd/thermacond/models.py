"""Synthetic crystals and their force constants, written in THERMACOND's input format.

The reference takes harmonic and cubic force constants from FOCEX fits
(FOCEX/ios.f90:1180-1260 writes fc2.dat/fc3.dat, :2256-2350 writes lat_fc.dat).  Here
they come from model potentials instead, by central finite differences of the forces:

    Phi2(ia,kb)    = -[F_kb(+h_ia) - F_kb(-h_ia)] / (2h)
    Phi3(ia,jb,kg) = -[F_kg(+h_ia,+h_jb) - F_kg(+h_ia,-h_jb)
                       - F_kg(-h_ia,+h_jb) + F_kg(-h_ia,-h_jb)] / (4h^2)

with forces F = -dE/du taken exactly (torch autograd, float64) on a finite cluster that
contains every atom interacting with the displaced ones.  Because the cluster forces sum
to zero for any displacement, both tensors obey the acoustic sum rule to rounding.

Models
  * Stillinger-Weber silicon (Stillinger & Weber, PRB 31, 5262 (1985)):
    eps=2.1683 eV, sigma=2.0951 A, A=7.049556277, B=0.6022245584, p=4, q=0, a=1.80,
    lambda=21.0, gamma=1.20; diamond lattice a=5.43 A, m=28.0855 amu.
  * Born-Mayer pair crystals (rock-salt and zincblende type): V(r) = A exp(-r/rho) for
    unlike neighbours with a harmonic bond-bending-free stabiliser; see born_mayer_*.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

# ---------------------------------------------------------------------------
# crystal description
# ---------------------------------------------------------------------------


@dataclasses.dataclass
class Crystal:
    """A crystal in the reference's conventions.

    latparams   a,b,c,alpha,beta,gamma of the conventional cell (a,b,c in units of
                `lattice_parameter`), as structure.params line 1 (rest.f90:96)
    primlatt    3x3, column j = j-th primitive vector in conventional units (rest.f90:97)
    lattice_parameter  scale in Angstrom (rest.f90:98)
    types       1-based type of each primitive atom
    masses      mass (amu) per type
    names       element label per type
    frac        (natoms0,3) positions in conventional reduced coordinates (rest.f90:120)
    """

    latparams: Tuple[float, ...]
    primlatt: np.ndarray
    lattice_parameter: float
    types: List[int]
    masses: List[float]
    names: List[str]
    frac: np.ndarray

    @property
    def natoms0(self) -> int:
        return len(self.types)

    @property
    def conv_to_cart(self) -> np.ndarray:
        """force_constants.f90:53-63 (angles in degrees)."""
        a, b, c, al, be, ga = self.latparams
        a, b, c = a * self.lattice_parameter, b * self.lattice_parameter, c * self.lattice_parameter
        d2r = math.atan(1.0) * 4 / 180.0
        al, be, ga = al * d2r, be * d2r, ga * d2r
        m = np.zeros((3, 3))
        m[0, 0] = a
        m[0, 1] = b * math.cos(ga)
        m[1, 1] = b * math.sin(ga)
        m[0, 2] = c * math.cos(be)
        m[1, 2] = c * (math.cos(al) - math.cos(ga) * math.cos(be)) / math.sin(ga)
        m[2, 2] = math.sqrt(c ** 2 - m[0, 2] ** 2 - m[1, 2] ** 2)
        return m

    @property
    def prim_cart(self) -> np.ndarray:
        """Rows are the Cartesian primitive vectors r1,r2,r3."""
        return (self.conv_to_cart @ self.primlatt).T

    def home_positions(self) -> np.ndarray:
        """Cartesian home-cell positions, folded into the primitive cell as
        force_constants.f90:148-156 (unitcell, force_constants.f90:1357-1377) does."""
        P = self.prim_cart.T  # columns = primitive vectors
        Pinv = np.linalg.inv(P)
        out = []
        for f in self.frac:
            x = self.conv_to_cart @ f
            b = Pinv @ x
            for i in range(3):
                while b[i] > 1.0 or abs(b[i] - 1) <= 1e-6:
                    b[i] -= 1
                while b[i] < 0.0 and abs(b[i]) > 1e-6:
                    b[i] += 1
            out.append(P @ b)
        return np.array(out)

    def atom_masses(self) -> np.ndarray:
        return np.array([self.masses[t - 1] for t in self.types], dtype=np.float64)


def diamond(a: float = 5.43, mass: float = 28.0855, name: str = "Si") -> Crystal:
    return Crystal((1, 1, 1, 90, 90, 90),
                   np.array([[0, .5, .5], [.5, 0, .5], [.5, .5, 0]], dtype=np.float64).T,
                   a, [1, 1], [mass], [name], np.array([[0, 0, 0], [.25, .25, .25]]))


def zincblende(a: float = 5.65, m1: float = 69.72, m2: float = 74.92,
               n1: str = "Ga", n2: str = "As") -> Crystal:
    return Crystal((1, 1, 1, 90, 90, 90),
                   np.array([[0, .5, .5], [.5, 0, .5], [.5, .5, 0]], dtype=np.float64).T,
                   a, [1, 2], [m1, m2], [n1, n2], np.array([[0, 0, 0], [.25, .25, .25]]))


def rocksalt_conventional(a: float = 5.6, m1: float = 72.63, m2: float = 78.97,
                          n1: str = "Ge", n2: str = "Se") -> Crystal:
    """8-atom simple-cubic cell of a rock-salt-type binary (4 + 4 atoms)."""
    fcc = [[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]]
    frac = [f for f in fcc] + [[(f[0] + .5) % 1, f[1], f[2]] for f in fcc]
    return Crystal((1, 1, 1, 90, 90, 90), np.eye(3), a, [1] * 4 + [2] * 4,
                   [m1, m2], [n1, n2], np.array(frac, dtype=np.float64))


def zincblende_tetragonal4(a: float = 5.65, m1: float = 69.72, m2: float = 74.92,
                           n1: str = "Ga", n2: str = "As") -> Crystal:
    """4-atom tetragonal cell of a zincblende-type binary: a1 = (1/2,-1/2,0) a,
    a2 = (1/2,1/2,0) a, a3 = (0,0,1) a (two primitive cells), 2 + 2 atoms."""
    prim = np.array([[.5, -.5, 0], [.5, .5, 0], [0, 0, 1]], dtype=np.float64).T
    frac = np.array([[0, 0, 0], [.5, 0, .5], [.25, .25, .25], [.75, .25, .75]], dtype=np.float64)
    return Crystal((1, 1, 1, 90, 90, 90), prim, a, [1, 1, 2, 2], [m1, m2], [n1, n2], frac)


def supercell_2x1x1(c: Crystal) -> Crystal:
    """2x1x1 supercell of a simple-cubic-cell crystal (primlatt = identity)."""
    assert np.allclose(c.primlatt, np.eye(3))
    frac = np.concatenate([c.frac, c.frac + np.array([1.0, 0, 0])]) * np.array([.5, 1, 1])
    lp = (2.0,) + tuple(c.latparams[1:])
    return Crystal(lp, np.eye(3), c.lattice_parameter, c.types + c.types,
                   c.masses, c.names, frac)


# ---------------------------------------------------------------------------
# potentials (torch, float64) and finite-difference force constants
# ---------------------------------------------------------------------------

SW_SI = dict(eps=2.1683, sigma=2.0951, A=7.049556277, B=0.6022245584, p=4.0, q=0.0,
             a=1.80, lam=21.0, gamma=1.20)


def _sw_energy(x, pairs_i, pairs_j, trip_c, trip_j, trip_k, P=SW_SI):
    import torch
    eps, sig, A, B, p, q, a, lam, gam = (P[k] for k in
                                         ("eps", "sigma", "A", "B", "p", "q", "a", "lam", "gamma"))
    rc = a * sig
    d = x[..., pairs_j, :] - x[..., pairs_i, :]
    r = torch.sqrt((d * d).sum(-1))
    e2 = eps * A * (B * (sig / r) ** p - (sig / r) ** q) * torch.exp(sig / (r - rc))
    dj = x[..., trip_j, :] - x[..., trip_c, :]
    dk = x[..., trip_k, :] - x[..., trip_c, :]
    rj = torch.sqrt((dj * dj).sum(-1))
    rk = torch.sqrt((dk * dk).sum(-1))
    cth = (dj * dk).sum(-1) / (rj * rk)
    e3 = eps * lam * torch.exp(gam * sig / (rj - rc) + gam * sig / (rk - rc)) * (cth + 1.0 / 3.0) ** 2
    return e2.sum() + e3.sum()


def _shell_energy(x, pi, pj, D, r0, delta, rho):
    """Born-Mayer-type double exponential per bonded pair (a Morse form with its minimum at
    the pair's equilibrium distance r0, so the lattice is in equilibrium and the harmonic
    force constants are radial springs k = 2D/rho^2), plus a cubic term that scales the
    third derivative -6D/rho^3 by (1 + delta)."""
    d = x[..., pj, :] - x[..., pi, :]
    r = (d * d).sum(-1).sqrt()
    u = (r - r0) / rho
    import torch
    return (D * (torch.exp(-2 * u) - 2 * torch.exp(-u)) - delta * D * u ** 3).sum()


def _lattice_cluster(crys: Crystal, radius: float):
    """All atoms with |R + tau_b - centre| < radius, home atoms first."""
    P = crys.prim_cart
    tau = crys.home_positions()
    centre = tau.mean(0)
    n = int(math.ceil(radius / min(np.linalg.norm(P, axis=1)))) + 2
    home, others = [], []
    for n1 in range(-n, n + 1):
        for n2 in range(-n, n + 1):
            for n3 in range(-n, n + 1):
                R = n1 * P[0] + n2 * P[1] + n3 * P[2]
                for b in range(crys.natoms0):
                    pos = R + tau[b]
                    if np.linalg.norm(pos - centre) < radius:
                        rec = (b + 1, (n1, n2, n3), pos)
                        (home if (n1, n2, n3) == (0, 0, 0) else others).append(rec)
    home.sort(key=lambda t: t[0])
    others.sort(key=lambda t: (round(float(np.linalg.norm(t[2] - centre)), 6), t[1], t[0]))
    return home + others


@dataclasses.dataclass
class ForceConstants:
    """Atom list (lat_fc.dat neighbour block) plus FC2/FC3 term lists.

    atompos (natoms,3) Cartesian; iatomcell0 (natoms,) 1-based basis atom;
    iatomcell (natoms,3) cell indices.  fc2: list of (i,al,j,be,value) with 1-based atom
    indices into the atom list, fc3: (i,al,j,be,k,ga,value)."""
    crystal: Crystal
    atompos: np.ndarray
    iatomcell0: np.ndarray
    iatomcell: np.ndarray
    fc2: np.ndarray     # (nt2, 5) float: i, al, j, be, value
    fc3: np.ndarray     # (nt3, 7) float: i, al, j, be, k, ga, value

    @property
    def natoms(self) -> int:
        return len(self.iatomcell0)


def _fd_force_constants(crys: Crystal, energy_fn, cutoff_pair: float, cutoff_fc3: float,
                        h: float = 0.01, radius: float = 10.0, rel_drop: float = 1e-9,
                        fc3_perturb=None, close_swap: bool = False, batched: bool = False) -> ForceConstants:
    """batched: evaluate all displaced clusters of one (i, alpha) in one autograd call (the
    energy functions take a leading batch axis); the values are the same as one call per
    configuration (False), only ten times faster."""
    import torch
    cl = _lattice_cluster(crys, radius)
    x0 = torch.tensor(np.array([c[2] for c in cl]), dtype=torch.float64)
    nat = x0.shape[0]

    def forces(x):
        x = x.clone().requires_grad_(True)
        e = energy_fn(x)
        (g,) = torch.autograd.grad(e, x)
        return (-g).detach()

    home = list(range(crys.natoms0))
    # FC2
    fc2 = {}
    for i in home:
        for al in range(3):
            xp = x0.clone(); xp[i, al] += h
            xm = x0.clone(); xm[i, al] -= h
            d = -(forces(xp) - forces(xm)) / (2 * h)
            for k in range(nat):
                for be in range(3):
                    v = float(d[k, be])
                    if v != 0.0:
                        fc2[(i, al, k, be)] = v
    # FC3: j within cutoff_fc3 of i
    fc3 = {}
    signs = ((1, 1), (1, -1), (-1, 1), (-1, -1))
    for i in home:
        near = [j for j in range(nat) if float(torch.linalg.norm(x0[j] - x0[i])) < cutoff_fc3]
        for al in range(3):
            if batched:
                X = x0.repeat(len(near) * 3 * 4, 1, 1)
                n = 0
                for j in near:
                    for be in range(3):
                        for si, sj in signs:
                            X[n, i, al] += si * h
                            X[n, j, be] += sj * h
                            n += 1
                FB = forces(X)
            n = 0
            for j in near:
                for be in range(3):
                    if batched:
                        f = [FB[n + t] for t in range(4)]
                        n += 4
                    else:
                        f = []
                        for si, sj in signs:
                            x = x0.clone(); x[i, al] += si * h; x[j, be] += sj * h
                            f.append(forces(x))
                    d = (-(f[0] - f[1] - f[2] + f[3]) / (4 * h * h)).numpy()
                    for k, ga in zip(*np.nonzero(d)):
                        fc3[(i, al, j, be, int(k), int(ga))] = float(d[k, ga])
    mx2 = max(abs(v) for v in fc2.values())
    mx3 = max(abs(v) for v in fc3.values())
    fc2 = {k: v for k, v in fc2.items() if abs(v) > rel_drop * mx2}
    # symmetrise FC3 in (j,be) <-> (k,ga) (finite-difference error is O(h^2)).  With
    # close_swap a term whose partner was not differentiated (k beyond cutoff_fc3, e.g.
    # Phi(i,i,k) of a pair potential) gives the partner its own value, so the list is closed
    # under the swap: the reference sizes its phase table by the largest second atom index
    # (cal_eiqr2, sub.f90:6035) and reads it with the third, which is only in bounds when the
    # third never exceeds the second's range.
    sym = {}
    for (i, al, j, be, k, ga), v in fc3.items():
        key2 = (i, al, k, ga, j, be)
        if close_swap:
            w = 0.5 * (v + fc3[key2]) if key2 in fc3 else v
            sym[key2] = w
        else:   # the SW-Si lists of the golden fixtures: partners within cutoff_fc3 only
            w = 0.5 * (v + fc3.get(key2, 0.0))
        sym[(i, al, j, be, k, ga)] = w
    fc3 = {k: v for k, v in sym.items()
           if abs(v) > rel_drop * mx3 or abs(sym.get((k[0], k[1], k[4], k[5], k[2], k[3]), 0.0)) > rel_drop * mx3}
    if fc3_perturb is not None:
        fc3 = fc3_perturb(fc3)
    used = set()
    for (i, al, k, be) in fc2:
        used.add(i); used.add(k)
    for (i, al, j, be, k, ga) in fc3:
        used.add(i); used.add(j); used.add(k)
    order = [a for a in range(nat) if a in used or a < crys.natoms0]
    remap = {a: n + 1 for n, a in enumerate(order)}
    atompos = np.array([cl[a][2] for a in order])
    iat0 = np.array([cl[a][0] for a in order], dtype=np.int64)
    iatc = np.array([cl[a][1] for a in order], dtype=np.int64)
    fc2a = np.array(sorted([(remap[i], al + 1, remap[k], be + 1, v) for (i, al, k, be), v in fc2.items()]),
                    dtype=np.float64)
    fc3a = np.array(sorted([(remap[i], al + 1, remap[j], be + 1, remap[k], ga + 1, v)
                            for (i, al, j, be, k, ga), v in fc3.items()]), dtype=np.float64)
    return ForceConstants(crys, atompos, iat0, iatc, fc2a, fc3a)


def sw_silicon(a: float = 5.43, h: float = 0.01, batched: bool = True) -> ForceConstants:
    """Stillinger-Weber Si force constants: FC2 to the 2nd shell, FC3 to the 2nd shell
    (the SW range 1.8*sigma = 3.77 A reaches only first neighbours, so three-body terms
    couple first and second neighbours)."""
    import torch
    crys = diamond(a)
    cl = _lattice_cluster(crys, 10.0)
    x0 = np.array([c[2] for c in cl])
    rc = SW_SI["a"] * SW_SI["sigma"]
    pi, pj, tc, tj, tk = [], [], [], [], []
    nbr = [[] for _ in range(len(x0))]
    for i in range(len(x0)):
        d = np.linalg.norm(x0 - x0[i], axis=1)
        for j in np.nonzero((d < rc) & (d > 0))[0]:
            nbr[i].append(int(j))
            if j > i:
                pi.append(i); pj.append(int(j))
    for c in range(len(x0)):
        for u in range(len(nbr[c])):
            for v in range(u + 1, len(nbr[c])):
                tc.append(c); tj.append(nbr[c][u]); tk.append(nbr[c][v])
    T = lambda l: torch.tensor(l, dtype=torch.long)
    pi, pj, tc, tj, tk = T(pi), T(pj), T(tc), T(tj), T(tk)
    fn = lambda x: _sw_energy(x, pi, pj, tc, tj, tk)
    nn = a * math.sqrt(3) / 4
    return _fd_force_constants(crys, fn, cutoff_pair=rc, cutoff_fc3=a / math.sqrt(2) * 1.05, h=h, batched=batched)


def _sw_force_constants(crys: Crystal, P: dict, a: float, h: float) -> ForceConstants:
    import torch
    cl = _lattice_cluster(crys, 10.0 + 0.5 * a * (crys.natoms0 // 2 - 1))
    x0 = np.array([c[2] for c in cl])
    rc = P["a"] * P["sigma"]
    pi, pj, tc, tj, tk = [], [], [], [], []
    nbr = [[] for _ in range(len(x0))]
    for i in range(len(x0)):
        d = np.linalg.norm(x0 - x0[i], axis=1)
        for j in np.nonzero((d < rc) & (d > 0))[0]:
            nbr[i].append(int(j))
            if j > i:
                pi.append(i); pj.append(int(j))
    for c in range(len(x0)):
        for u in range(len(nbr[c])):
            for v in range(u + 1, len(nbr[c])):
                tc.append(c); tj.append(nbr[c][u]); tk.append(nbr[c][v])
    T = lambda l: torch.tensor(l, dtype=torch.long)
    pi, pj, tc, tj, tk = T(pi), T(pj), T(tc), T(tj), T(tk)
    fn = lambda x: _sw_energy(x, pi, pj, tc, tj, tk, P)
    return _fd_force_constants(crys, fn, cutoff_pair=rc, cutoff_fc3=a / math.sqrt(2) * 1.05, h=h,
                               radius=10.0 + 0.5 * a * (crys.natoms0 // 2 - 1), close_swap=True, batched=True)


def sw_zincblende4(a: float = 5.65, h: float = 0.01) -> ForceConstants:
    """4-atom zincblende-type cell (masses 69.72/74.92) with the Stillinger-Weber form, its
    length scale sigma stretched by a/5.43 so the lattice sits at the same reduced geometry
    as SW silicon (a tetrahedral bond-angle term keeps the zincblende structure stable)."""
    P = dict(SW_SI)
    P["sigma"] = SW_SI["sigma"] * a / 5.43
    return _sw_force_constants(zincblende_tetragonal4(a), P, a, h)


# shells of the rock-salt structure: (distance / a, species relation, spring constant eV/A^2)
ROCKSALT_SHELLS = ((0.5, "unlike", 2.5), (math.sqrt(0.5), "like", 0.5), (math.sqrt(0.75), "unlike", 0.12))


def born_mayer_rocksalt(crys: Optional[Crystal] = None, a: float = 5.6, rho: float = 0.35, seed: int = 0,
                        perturb: float = 0.05, h: float = 0.01) -> ForceConstants:
    """Rock-salt-type binary with Born-Mayer-type pair springs (no Coulomb term).

    Harmonic force constants reach the 3rd shell (unlike pairs at a/2, like pairs at a/sqrt2,
    unlike pairs at a sqrt3/2; ROCKSALT_SHELLS).  The cubic force constants of a pair
    potential couple a home atom to its bonded partners only, and they are kept for every
    bonded shell: truncating them at the 2nd shell would leave Phi(i,i,k) of the 3rd-shell
    bonds without their swap partners Phi(i,k,i), Phi(i,k,k), breaking the (j,k) symmetry and
    the acoustic sum rule.  The cubic part carries a seeded Gaussian perturbation of relative width
    `perturb` (5 %): one draw per symmetry class of bonds (shell x species pair), applied to
    the pair potential's third derivative, so the perturbed constants keep the space-group
    symmetry and the acoustic sum rule exactly.  `crys` defaults to the 8-atom cubic cell;
    a supercell of it (supercell_2x1x1) gives the 16-atom cell."""
    import torch
    crys = rocksalt_conventional(a) if crys is None else crys
    radius = 10.5 if crys.natoms0 <= 8 else 12.5
    cl = _lattice_cluster(crys, radius)
    x0 = np.array([c[2] for c in cl])
    ty = np.array([crys.types[c[0] - 1] for c in cl])
    rng = np.random.default_rng(seed)
    classes = {}
    pi, pj, D, r0, dl = [], [], [], [], []
    for i in range(len(x0)):
        d = np.linalg.norm(x0 - x0[i], axis=1)
        for j in range(i + 1, len(x0)):
            for s, (f, rel, k) in enumerate(ROCKSALT_SHELLS):
                if abs(d[j] - f * a) < 1e-3 and ((ty[i] != ty[j]) == (rel == "unlike")):
                    key = (s, min(ty[i], ty[j]), max(ty[i], ty[j]))
                    if key not in classes:
                        classes[key] = float(rng.standard_normal()) * perturb
                    pi.append(i); pj.append(j); D.append(k * rho * rho / 2); r0.append(f * a)
                    dl.append(classes[key])
    T = lambda l, dt: torch.tensor(l, dtype=dt)
    pi, pj = T(pi, torch.long), T(pj, torch.long)
    D, r0, dl = T(D, torch.float64), T(r0, torch.float64), T(dl, torch.float64)
    fn = lambda x: _shell_energy(x, pi, pj, D, r0, dl, rho)
    return _fd_force_constants(crys, fn, cutoff_pair=ROCKSALT_SHELLS[-1][0] * a * 1.01,
                               cutoff_fc3=ROCKSALT_SHELLS[-1][0] * a * 1.05, h=h, radius=radius, close_swap=True,
                               batched=True)


def born_mayer_supercell16(a: float = 5.6, seed: int = 0, h: float = 0.01) -> ForceConstants:
    """16-atom 2x1x1 supercell of the 8-atom rock-salt-type cell (same potential)."""
    return born_mayer_rocksalt(supercell_2x1x1(rocksalt_conventional(a)), a=a, seed=seed, h=h)

