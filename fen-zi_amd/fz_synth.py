"""Synthetic MD workloads: rigid 3-site waters and linear 38-bead chains.

Builds the five benchmark systems (water box; 17K-atom slab; 230K-atom slab;
batched ensembles; serving jobs) as plain numpy topology arrays in the layout of
``fz_topology`` (include/fenzi.h), plus coordinates and run parameters.  The
same arrays can be written out as a FEN ZI input deck (.in/.psf/.prm/.crd) with
``write_deck`` so that the parser path is exercised too.

Parameters follow the workload definitions: TIP3P-like water (O -0.834,
H +0.417, LJ on O only, eps 0.1521, Rmin/2 1.7682; OH 0.9572 A, HOH 104.52 deg,
rigid by SHAKE); chain beads eps 0.08, Rmin/2 2.0, bonds k=300 r0=1.53,
angles k=58 theta0=113.6, dihedrals k=0.2 n=3 delta=0; head dipole +-0.5 on
beads 1-2.  Bead mass 14.027 amu (united-atom CH2).
"""
from __future__ import annotations

import math

import numpy as np

# Atom types: 0 = water O (OT), 1 = water H (HT), 2 = chain bead (CB)
TYPE_NAMES = ["OT", "HT", "CB"]
TYPE_MASS = np.array([15.9994, 1.008, 14.027])
TYPE_EPS = np.array([0.1521, 0.0, 0.08])       # kcal/mol (positive well depth)
TYPE_RMIN2 = np.array([1.7682, 0.0, 2.0])      # Rmin/2, A

Q_O, Q_H = -0.834, 0.417
R_OH, ANG_HOH = 0.9572, math.radians(104.52)
R_HH = 2.0 * R_OH * math.sin(ANG_HOH / 2.0)    # 1.5139 A
CHAIN_LEN = 38
B_K, B_R0 = 300.0, 1.53
A_K, A_T0 = 58.0, math.radians(113.6)
D_K, D_N, D_DELTA = 0.2, 3, 0.0
WATER_BOND_K = 450.0
WATER_ANGLE_K = 55.0
AMU_A3_TO_GCC = 1.66053907


def lj_tables(types_present=3):
    """Pair tables eps_ij = 4 sqrt(eps_i eps_j), sigma_ij = sigma_i/2 + sigma_j/2,
    sigma/2 = (Rmin/2) / 2^(1/6) (reference src/prmread.cu:2468-2471,2725-2735)."""
    e = TYPE_EPS[:types_present]
    s_half = TYPE_RMIN2[:types_present] / 2.0 ** (1.0 / 6.0)
    eps = 4.0 * np.sqrt(np.outer(e, e))
    sig = s_half[:, None] + s_half[None, :]
    return eps, sig * sig


def smooth235(n):
    while True:
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            return n
        n += 1


def random_rotations(rng, n):
    """Uniform random rotation matrices (Shoemake's quaternion method)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    q = np.stack([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                  b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)], axis=1)
    x, y, z, w = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def water_local():
    h = ANG_HOH / 2.0
    return np.array([[0.0, 0.0, 0.0],
                     [R_OH * math.sin(h), 0.0, R_OH * math.cos(h)],
                     [-R_OH * math.sin(h), 0.0, R_OH * math.cos(h)]])


def _place_waters(rng, sites, jitter):
    n = len(sites)
    R = random_rotations(rng, n)
    loc = water_local()
    o = sites + rng.uniform(-jitter, jitter, size=(n, 3))
    xyz = o[:, None, :] + np.einsum("nij,kj->nki", R, loc)
    return xyz.reshape(-1, 3)


def _lattice_sites(box, zlo, zhi, n_needed, spacing=None):
    """Cubic-ish lattice filling x in [-L/2, L/2), y likewise, z in [zlo, zhi)."""
    Lx, Ly = box[0], box[1]
    H = zhi - zlo
    if spacing is None:
        spacing = (Lx * Ly * H / n_needed) ** (1.0 / 3.0)
    a = spacing
    while True:
        nx, ny, nz = max(1, int(Lx / a)), max(1, int(Ly / a)), max(1, int(H / a))
        if nx * ny * nz >= n_needed:
            break
        a *= 0.98
    ax, ay, az = Lx / nx, Ly / ny, H / nz
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pts = np.stack([-Lx / 2 + (ix.ravel() + 0.5) * ax,
                    -Ly / 2 + (iy.ravel() + 0.5) * ay,
                    zlo + (iz.ravel() + 0.5) * az], axis=1)
    return pts


def _grid_dims(n, Lx, Ly):
    best = None
    for a in range(1, n + 1):
        if n % a:
            continue
        b = n // a
        score = abs(math.log((a / b) / (Lx / Ly)))
        if best is None or score < best[0]:
            best = (score, a, b)
    if best[0] > 0.7:   # awkward factorisation: use a near-square grid and keep the first n sites
        a = int(math.ceil(math.sqrt(n * Lx / Ly)))
        b = int(math.ceil(n / a))
        return a, b
    return best[1], best[2]


def _grow_chains(rng, heads, axes, direction, zlo, zhi, radius):
    """Grow linear chains bead by bead from head positions: fixed bond length and
    bond angle, dihedral picked among 18 candidates so that the bead stays within
    `radius` of its column axis and inside [zlo, zhi], and away from all of its
    own earlier beads."""
    m = len(heads)
    pos = np.zeros((m, CHAIN_LEN, 3))
    pos[:, 0] = heads
    pos[:, 1] = heads + np.array([0, 0, direction * B_R0])
    # third bead: in a random azimuthal plane
    phi = rng.uniform(0, 2 * np.pi, m)
    ang = np.pi - A_T0
    step = np.stack([np.sin(ang) * np.cos(phi), np.sin(ang) * np.sin(phi),
                     np.full(m, direction * np.cos(ang))], axis=1) * B_R0
    pos[:, 2] = pos[:, 1] + step
    base = np.radians(np.arange(-180.0, 180.0, 20.0))
    for k in range(3, CHAIN_LEN):
        a, b, c = pos[:, k - 3], pos[:, k - 2], pos[:, k - 1]
        bc = c - b
        bc /= np.linalg.norm(bc, axis=1)[:, None]
        n = np.cross(b - a, bc)
        n /= np.linalg.norm(n, axis=1)[:, None] + 1e-300
        mvec = np.cross(n, bc)
        cands = []
        for d in base:
            dd = d + rng.uniform(-0.25, 0.25, m)
            # NeRF placement of a bead given bond, angle, dihedral
            x = -B_R0 * math.cos(A_T0)
            y = B_R0 * math.sin(A_T0) * np.cos(dd)
            z = B_R0 * math.sin(A_T0) * np.sin(dd)
            cands.append(c + x * bc + y[:, None] * mvec + z[:, None] * n)
        cands = np.stack(cands, axis=1)                     # (m, ncand, 3)
        dxy = np.sqrt(((cands[:, :, :2] - axes[:, None, :]) ** 2).sum(-1))
        score = 20.0 * np.maximum(dxy - radius, 0.0) ** 2
        over = np.maximum(zlo - cands[:, :, 2], 0) + np.maximum(cands[:, :, 2] - zhi, 0)
        score += 25.0 * over ** 2
        # drift towards the slab centre at the rate that spreads 38 beads over the leaflet
        zt = heads[:, 2] + direction * (k / (CHAIN_LEN - 1.0)) * (zhi - zlo)
        score += 0.5 * (cands[:, :, 2] - zt[:, None]) ** 2
        if k >= 4:
            prev = pos[:, : k - 3]
            dist = np.linalg.norm(cands[:, :, None, :] - prev[:, None, :, :], axis=-1)
            score += 10.0 * (np.maximum(3.4 - dist, 0.0) ** 2).sum(-1)
        score += rng.uniform(0, 0.3, score.shape)
        pick = np.argmin(score, axis=1)
        pos[:, k] = cands[np.arange(m), pick]
    return pos


def _relax_beads(x, box, dmin=2.6, iters=80):
    """Push apart bead pairs closer than dmin that are more than 3 bonds apart
    (periodic KD-tree), then restore the bond lengths along each chain."""
    from scipy.spatial import cKDTree
    x = x.copy()
    n = len(x)
    for _ in range(iters):
        w = np.mod(x + box / 2, box)
        w = np.minimum(w, np.nextafter(box, 0))
        pairs = cKDTree(w, boxsize=box).query_pairs(dmin, output_type="ndarray")
        if len(pairs):
            i, j = pairs[:, 0], pairs[:, 1]
            keep = ~((i // CHAIN_LEN == j // CHAIN_LEN) & (np.abs(i - j) <= 3))
            i, j = i[keep], j[keep]
        if len(pairs) == 0 or len(i) == 0:
            break
        d = x[j] - x[i]
        d -= box * np.round(d / box)
        r = np.maximum(np.linalg.norm(d, axis=1), 1e-6)
        push = (0.3 * (dmin - r) / r)[:, None] * d
        delta = np.zeros_like(x)
        np.add.at(delta, i, -push)
        np.add.at(delta, j, push)
        x += delta
        xc = x.reshape(-1, CHAIN_LEN, 3)
        for _sweep in range(6):
            for par in (0, 1):
                a = xc[:, par:CHAIN_LEN - 1:2]
                b = xc[:, par + 1:CHAIN_LEN:2]
                m = min(a.shape[1], b.shape[1])
                a, b = a[:, :m], b[:, :m]
                v = b - a
                L = np.linalg.norm(v, axis=-1, keepdims=True)
                corr = 0.5 * (L - B_R0) / L * v
                xc[:, par:par + 2 * m:2] += corr
                xc[:, par + 1:par + 1 + 2 * m:2] -= corr
        x = xc.reshape(n, 3)
    return x


def make_system(n_chains, n_waters, box, seed=0, jitter=0.3, water_grid=None,
                shake=True, leaflets=2):
    """Topology + coordinates for chains (two leaflets of a slab centred at z=0)
    followed by rigid waters.  Returns a dict of numpy arrays."""
    rng = np.random.default_rng(seed)
    box = np.asarray(box, dtype=np.float64)
    Lx, Ly, Lz = box
    n_beads = n_chains * CHAIN_LEN
    n_atoms = n_beads + 3 * n_waters
    xyz_parts = []
    h = 0.0
    if n_chains:
        # slab thickness from the natural volumes of a CH2 bead (27 A^3) and a water (30 A^3)
        h = 0.5 * Lz * (27.0 * n_beads) / (27.0 * n_beads + 30.0 * n_waters)
        h = min(max(h, 5.0), 0.45 * Lz)  # room for a few chains in a mostly-water box
        per = [n_chains // 2 + n_chains % 2, n_chains // 2] if leaflets == 2 else [n_chains]
        for li, m in enumerate(per):
            if m == 0:
                continue
            gx, gy = _grid_dims(m, Lx, Ly)
            ix, iy = np.meshgrid(np.arange(gx), np.arange(gy), indexing="xy")
            ax = np.stack([-Lx / 2 + (ix.ravel() + 0.5) * Lx / gx,
                           -Ly / 2 + (iy.ravel() + 0.5) * Ly / gy], axis=1)[:m]
            ax = ax + rng.uniform(-0.5, 0.5, size=ax.shape)
            upper = (li == 0)
            zhead = (h - 1.0) if upper else (-h + 1.0)
            heads = np.column_stack([ax, np.full(m, zhead)])
            zlo, zhi = (0.3, h - 0.5) if upper else (-h + 0.5, -0.3)
            radius = 0.42 * min(Lx / gx, Ly / gy)
            pos = _grow_chains(rng, heads, ax, -1.0 if upper else 1.0, zlo, zhi, radius)
            xyz_parts.append(pos.reshape(-1, 3))
        xyz_parts = [_relax_beads(np.concatenate(xyz_parts), box)]
    if n_waters:
        if water_grid is not None:
            nside, spacing = water_grid
            idx = np.arange(n_waters)
            ix, iy, iz = idx % nside, (idx // nside) % nside, idx // (nside * nside)
            sites = np.stack([ix, iy, iz], axis=1) * spacing + spacing / 2.0 - box / 2.0
        elif n_chains:
            # water above and below the slab: a lattice over the wrapped range [h, Lz - h),
            # without the sites that fall within 2.8 A of a bead
            from scipy.spatial import cKDTree
            beads = np.mod(xyz_parts[0] + box / 2, box)
            beads = np.minimum(beads, np.nextafter(box, 0))
            tree = cKDTree(beads, boxsize=box)
            want = n_waters
            while True:
                pts = _lattice_sites(box, h + 1.0, Lz - h - 1.0, want)
                pts[:, 2] = (pts[:, 2] + Lz / 2) % Lz - Lz / 2
                w = np.minimum(np.mod(pts + box / 2, box), np.nextafter(box, 0))
                d, _ = tree.query(w)
                pts = pts[d > 2.8]
                if len(pts) >= n_waters:
                    break
                want = int(want * 1.1) + 8
            sel = rng.permutation(len(pts))[:n_waters]
            sites = pts[np.sort(sel)]
        else:
            pts = _lattice_sites(box, -Lz / 2, Lz / 2, n_waters)
            sel = rng.permutation(len(pts))[:n_waters]
            sites = pts[np.sort(sel)]
        xyz_parts.append(_place_waters(rng, sites, jitter))
    xyz = np.concatenate(xyz_parts, axis=0) if xyz_parts else np.zeros((0, 3))
    xyz = (xyz + box / 2) % box - box / 2

    ntypes = 3 if n_chains else 2
    typ = np.empty(n_atoms, np.int32)
    q = np.zeros(n_atoms)
    typ[:n_beads] = 2
    if n_chains:
        q[0:n_beads:CHAIN_LEN] = 0.5
        q[1:n_beads:CHAIN_LEN] = -0.5
    w0 = n_beads
    typ[w0::3] = 0; typ[w0 + 1::3] = 1; typ[w0 + 2::3] = 1
    q[w0::3] = Q_O; q[w0 + 1::3] = Q_H; q[w0 + 2::3] = Q_H
    mass = TYPE_MASS[typ]

    # chains
    cb = np.arange(n_chains)[:, None] * CHAIN_LEN
    b_ch = (cb + np.arange(CHAIN_LEN - 1)[None, :]).reshape(-1)
    bonds_c = np.stack([b_ch, b_ch + 1], 1)
    a_ch = (cb + np.arange(CHAIN_LEN - 2)[None, :]).reshape(-1)
    angles_c = np.stack([a_ch, a_ch + 1, a_ch + 2], 1)
    d_ch = (cb + np.arange(CHAIN_LEN - 3)[None, :]).reshape(-1)
    dihed_c = np.stack([d_ch, d_ch + 1, d_ch + 2, d_ch + 3], 1)
    # waters
    wo = w0 + 3 * np.arange(n_waters)
    bonds_w = np.concatenate([np.stack([wo, wo + 1], 1), np.stack([wo, wo + 2], 1),
                              np.stack([wo + 1, wo + 2], 1)])
    angles_w = np.stack([wo + 1, wo, wo + 2], 1)
    bond_atoms = np.concatenate([bonds_c, bonds_w]).astype(np.int32)
    bond_k = np.concatenate([np.full(len(bonds_c), B_K), np.full(2 * n_waters, WATER_BOND_K),
                             np.zeros(n_waters)])
    bond_r0 = np.concatenate([np.full(len(bonds_c), B_R0), np.full(2 * n_waters, R_OH),
                              np.full(n_waters, R_HH)])
    angle_atoms = np.concatenate([angles_c, angles_w]).astype(np.int32)
    angle_k = np.concatenate([np.full(len(angles_c), A_K), np.full(n_waters, WATER_ANGLE_K)])
    angle_t0 = np.concatenate([np.full(len(angles_c), A_T0), np.full(n_waters, ANG_HOH)])
    eps, sig2 = lj_tables(ntypes)
    topo = dict(
        n_atoms=n_atoms, n_types=ntypes, charge=q, mass=mass, type=typ,
        lj_eps=eps.ravel().copy(), lj_sig2=sig2.ravel().copy(),
        lj14_eps=eps.ravel().copy(), lj14_sig2=sig2.ravel().copy(),
        bond_atoms=bond_atoms.ravel(), bond_k=bond_k, bond_r0=bond_r0,
        angle_atoms=angle_atoms.ravel(), angle_k=angle_k, angle_theta0=angle_t0,
        ub_k=np.zeros(len(angle_atoms)), ub_s0=np.zeros(len(angle_atoms)),
        dihedral_atoms=dihed_c.astype(np.int32).ravel(), dihedral_k=np.full(len(dihed_c), D_K),
        dihedral_n=np.full(len(dihed_c), D_N, np.int32), dihedral_delta=np.full(len(dihed_c), D_DELTA),
        improper_atoms=np.zeros(0, np.int32), improper_k=np.zeros(0), improper_psi0=np.zeros(0),
    )
    if shake:
        light = mass[bond_atoms[:, 0]] <= 1.5
        light |= mass[bond_atoms[:, 1]] <= 1.5
        topo["constraint_atoms"] = bond_atoms[light].astype(np.int32).ravel()
        topo["constraint_length"] = bond_r0[light].copy()
    else:
        topo["constraint_atoms"] = np.zeros(0, np.int32)
        topo["constraint_length"] = np.zeros(0)
    for k in ("bond", "angle", "dihedral", "improper", "constraint"):
        topo["n_%ss" % k] = len(topo["%s_atoms" % k]) // (2 if k in ("bond", "constraint") else 3 if k == "angle" else 4)
    return topo, xyz


def config(cid, seed=None, **over):
    """Return (topo, xyz, params) for workload cid in {0, 1, 2} or a dict for 3/4."""
    seed = cid if seed is None else seed
    if cid == 0:
        box = (30.0, 30.0, 30.0)
        topo, xyz = make_system(0, 2400, box, seed=seed, jitter=0.3, water_grid=(14, 30.0 / 14))
        params = dict(box=box, cutoff=10.0, cuton=8.0, pairlist=11.5, beta=0.34, K=(32, 32, 32),
                      dt_ps=0.001, shake_tol=1e-8, temperature=300.0, steps=1000, ensemble="NVE")
    elif cid == 1:
        box = (48.0, 48.0, 72.0)
        topo, xyz = make_system(128, 4046, box, seed=seed)
        params = dict(box=box, cutoff=12.0, cuton=10.0, pairlist=13.5, beta=0.26, K=(48, 48, 72),
                      dt_ps=0.002, shake_tol=1e-8, temperature=300.0, steps=5000, ensemble="NVT",
                      reassign_every=500)
    elif cid == 2:
        box = (132.0, 132.0, 132.0)
        topo, xyz = make_system(1536, 57216, box, seed=seed)
        params = dict(box=box, cutoff=12.0, cuton=10.0, pairlist=13.5, beta=0.26, K=(128, 128, 128),
                      dt_ps=0.002, shake_tol=1e-8, temperature=300.0, steps=2000, ensemble="NVT",
                      reassign_every=500)
    else:
        raise ValueError("use ensemble_member() for workloads 3 and 4")
    params.update(over)
    return topo, xyz, params


def ensemble_member(n_target, chain_frac, seed, cutoff=10.0):
    """One periodic water-plus-chain system of about n_target atoms in a cubic box
    at ~1 g/cm^3 with a PME grid near 1 A (2,3,5-smooth, clamped to 16..48)."""
    n_chains = int(round(chain_frac * n_target / CHAIN_LEN))
    n_waters = max(1, int(round((n_target - CHAIN_LEN * n_chains) / 3)))
    mass = n_chains * CHAIN_LEN * TYPE_MASS[2] + n_waters * (TYPE_MASS[0] + 2 * TYPE_MASS[1])
    L = (mass * AMU_A3_TO_GCC / 1.0) ** (1.0 / 3.0)
    K = min(48, max(16, smooth235(int(math.ceil(L)))))
    topo, xyz = make_system(n_chains, n_waters, (L, L, L), seed=seed)
    skin = min(1.5, max(0.2, L / 2 - cutoff - 0.05))
    params = dict(box=(L, L, L), cutoff=cutoff, cuton=cutoff - 2.0, pairlist=cutoff + skin,
                  beta=0.3125, K=(K, K, K), dt_ps=0.002, shake_tol=1e-8, temperature=300.0)
    return topo, xyz, params


def write_deck(topo, xyz, params, directory, name="synth", temperature=300.0, steps=1000, shake=True,
               reassign_every=0, output_every=100, seed=1, psf_seg=None, crd_seg=None, extra=()):
    """Write a FEN ZI input deck (.in / .psf / .prm / .crd) for a system built
    by make_system, with string atom types (no .rtf), in CHARMM formats.
    psf_seg / crd_seg: per-atom segment names of the PSF and CRD files (default
    "SYS"); extra: further keyword lines of the .in file."""
    import os
    os.makedirs(directory, exist_ok=True)
    n = int(topo["n_atoms"])
    tnames = TYPE_NAMES[: int(topo["n_types"])]
    res = ["CHN" if t == 2 else "TIP3" for t in topo["type"]]
    with open(os.path.join(directory, name + ".psf"), "w") as f:
        f.write("PSF\n\n       1 !NTITLE\n* synthetic system\n\n")
        f.write("%8d !NATOM\n" % n)
        for i in range(n):
            t = int(topo["type"][i])
            f.write("%8d %-4s %-4d %-4s %-4s %-4s %14.6f %14.6f %8d\n" % (
                i + 1, psf_seg[i] if psf_seg is not None else "SYS", i // 3 + 1, res[i], tnames[t], tnames[t],
                topo["charge"][i], topo["mass"][i], 0))

        def section(tag, arr, width, per_line):
            arr = np.asarray(arr).reshape(-1, width) + 1
            f.write("\n%8d %s\n" % (len(arr), tag))
            for k in range(0, len(arr), per_line):
                f.write("".join("%8d" % v for v in arr[k:k + per_line].ravel()) + "\n")
        section("!NBOND: bonds", topo["bond_atoms"], 2, 4)
        section("!NTHETA: angles", topo["angle_atoms"], 3, 3)
        section("!NPHI: dihedrals", topo["dihedral_atoms"], 4, 2)
        section("!NIMPHI: impropers", topo["improper_atoms"], 4, 2)
    with open(os.path.join(directory, name + ".prm"), "w") as f:
        f.write("* synthetic parameters\n*\n\nBONDS\n")
        seen = set()
        ba = np.asarray(topo["bond_atoms"]).reshape(-1, 2)
        for (a, b), k, r0 in zip(ba, topo["bond_k"], topo["bond_r0"]):
            key = (tnames[topo["type"][a]], tnames[topo["type"][b]])
            if key not in seen:
                seen.add(key)
                f.write("%-6s %-6s %16.10f %16.10f\n" % (key + (k, r0)))
        f.write("\nANGLES\n")
        seen = set()
        aa = np.asarray(topo["angle_atoms"]).reshape(-1, 3)
        for (a, b, c), k, t0 in zip(aa, topo["angle_k"], topo["angle_theta0"]):
            key = tuple(tnames[topo["type"][x]] for x in (a, b, c))
            if key not in seen:
                seen.add(key)
                f.write("%-6s %-6s %-6s %16.10f %18.12f\n" % (key + (k, math.degrees(t0))))
        f.write("\nDIHEDRALS\n")
        seen = set()
        da = np.asarray(topo["dihedral_atoms"]).reshape(-1, 4)
        for q, k, nn, dl in zip(da, topo["dihedral_k"], topo["dihedral_n"], topo["dihedral_delta"]):
            key = tuple(tnames[topo["type"][x]] for x in q) + (int(nn),)
            if key not in seen:
                seen.add(key)
                f.write("%-6s %-6s %-6s %-6s %16.10f %3d %16.10f\n" % (key[:4] + (k, nn, math.degrees(dl))))
        f.write("\nIMPROPER\n\nNONBONDED nbxmod  5 atom cdiel shift vatom vdistance vswitch -\n"
                "cutnb 14.0 ctofnb 12.0 ctonnb 10.0 eps 1.0 e14fac 1.0 wmin 1.5\n")
        for t, nm in enumerate(tnames):
            f.write("%-6s %10.6f %16.10f %16.10f\n" % (nm, 0.0, -TYPE_EPS[t], TYPE_RMIN2[t]))
        f.write("\nEND\n")
    with open(os.path.join(directory, name + ".crd"), "w") as f:
        f.write("* synthetic coordinates\n*\n%10d  EXT\n" % n)
        for i in range(n):
            t = int(topo["type"][i])
            f.write("%10d%10d  %-8s  %-8s%20.10f%20.10f%20.10f  %-8s  %-8d%20.10f\n" % (
                i + 1, i // 3 + 1, res[i], tnames[t], xyz[i, 0], xyz[i, 1], xyz[i, 2],
                crd_seg[i] if crd_seg is not None else "SYS", i // 3 + 1, 0.0))
    box = params["box"]
    K = params["K"]
    with open(os.path.join(directory, name + ".in"), "w") as f:
        f.write("# synthetic deck\ncoordinates %s.crd\nstructure %s.psf\nparameters %s.prm\n" % (name, name, name))
        f.write("temperature %g\nseed %d\ntimestep %g\ndynasteps %d\noutputTiming %d\n" % (
            temperature, seed, params["dt_ps"], steps, output_every))
        if reassign_every:
            f.write("tempresetTiming %d\n" % reassign_every)
        f.write("cuton %g\ncutoff %g\npairlistdist %g\nKAPPa %g\n" % (
            params["cuton"], params["cutoff"], params["pairlist"], params["beta"]))
        f.write("cellBasisVector1 %.10f 0 0\ncellBasisVector2 0 %.10f 0\ncellBasisVector3 0 0 %.10f\n" % tuple(box))
        f.write("PMEGridSizeX %d\nPMEGridSizeY %d\nPMEGridSizeZ %d\n" % tuple(K))
        f.write("shake %s %g\nprinttostdout 1\n" % ("on" if shake else "off", params.get("shake_tol", 1e-8)))
        for line in extra:
            f.write(line + "\n")
    return os.path.join(directory, name + ".in")


def batch_sizes(n_systems=256, lo=1500, hi=24000, seed=3):
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(math.log(lo), math.log(hi), n_systems)).astype(int)
    frac = rng.uniform(0.0, 0.3, n_systems)
    return n, frac


def serving_jobs(n_jobs=100, rate=20.0, lo=1000, hi=60000, seed=4):
    """Job stream of the serving workload: Poisson arrivals at `rate` jobs/s,
    N ~ LogUniform[lo, hi], chain fraction ~ U[0, 0.3], NVT length U{500..5000}."""
    rng = np.random.default_rng(seed)
    arrival = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
    n = np.exp(rng.uniform(math.log(lo), math.log(hi), n_jobs)).astype(int)
    frac = rng.uniform(0.0, 0.3, n_jobs)
    md_steps = rng.integers(500, 5001, n_jobs)
    return arrival, n, frac, md_steps


def _member_job(spec):
    n, frac, seed, cutoff = spec
    return ensemble_member(int(n), float(frac), int(seed), cutoff)


def ensemble_members(specs, workers=None):
    """ensemble_member over (n_target, chain_frac, seed, cutoff) specs, in worker
    processes (forked: call this before the process opens a GPU context)."""
    import multiprocessing as mp
    import os
    workers = workers or min(16, os.cpu_count() or 1)
    if workers <= 1 or len(specs) < 4:
        return [_member_job(s) for s in specs]
    with mp.get_context("fork").Pool(workers) as pool:
        return pool.map(_member_job, specs, chunksize=1)
