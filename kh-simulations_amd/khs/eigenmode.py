"""Linear SU(2) Kelvin-Helmholtz eigenmodes of the frozen log-cosh background,
and the eigen-seed files that mode-6 runs start from.

Restates ymgpu2d/analysis/ym_eigenmode.py:
  * wkb_growth_rate (:60-86): Im of the fastest-growing root of the WKB
    quartic  w^4 - kz^2 w^2 - (2n+1) sqrt(alpha^3/2) V0 w - alpha^2 V0 kz = 0;
  * build_operator (:89-198): the 6N x 6N complex operator L of
    gamma psi = L psi, psi = [b, ex, ez, a, qA, qB] (colour 2 + i colour 3),
    periodic in x with the simulation's stencils (forward d/dx ez in
    Faraday, backward d/dx b in Ampere), optional sponge -sigma_s (|xi| -
    xi_sponge) on every block and Dirichlet cut rows for |xi| > xi_cut;
  * solve (:201-234): the n_eigs eigenvalues nearest the shift sigma
    (default 0.55 gamma_WKB, else 0.1) by shift-invert implicitly restarted
    Arnoldi (scipy's ARPACK binding, which='LM', tol 1e-10), sorted by
    decreasing real part;
  * is_localised (:245-260) and growing_mode (the selection of :345-369 and
    presentation/compute_eigensolver_grid.py:58-67): the first eigenvalue
    with Re > 1e-5 whose |b| peaks inside 1.15 xi_inner and is below
    frac x peak at both ends;
  * seed_profiles / write_seed / seed_filename (:411-444): the six profiles
    phase-aligned so that Az peaks real and positive, scaled by max |Az|,
    real parts as float32 behind an int32 header {6, NX}; the file name
    carries kz, alpha, V0 and the sponge, plus EPS / Lz / Lx tags when they
    differ from 0.15 / 2 pi / 6 pi.

The operator is assembled with array operations instead of the reference's
per-row loop; the entries and their summation (sponge added to existing
diagonals) are the same.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

LX_DEFAULT = 6.0 * np.pi
LZ_DEFAULT = 2.0 * np.pi
EPS_DEFAULT = 0.15


def wkb_growth_rate(kz, alpha, V0, n_wkb=0):
    """gamma = max Im(w) over the WKB quartic's roots (NaN if none grows)."""
    if alpha == 0:
        return float("nan")
    C = (2 * n_wkb + 1) * np.sqrt(alpha ** 3 / 2.0) * V0
    roots = np.roots([1.0, 0.0, -kz ** 2, -C, -alpha ** 2 * V0 * kz])
    growing = roots[roots.imag > 1e-10]
    return float(np.max(growing.imag)) if len(growing) else float("nan")


def grid(NX, EPS, Lx=LX_DEFAULT):
    """x nodes and xi = (x - Lx/2) / EPS."""
    x = np.arange(NX) * (Lx / NX)
    return x, (x - Lx / 2.0) / EPS


def build_operator(kz, alpha, V0, EPS, NX, Lx=LX_DEFAULT, xi_sponge=None, sigma_sponge=5.0, xi_cut=None):
    """CSR matrix L (6N x 6N complex); state blocks b, ex, ez, a, qA, qB."""
    N = NX
    dx = Lx / N
    _, xi = grid(N, EPS, Lx)
    Az1 = -V0 * np.log(np.cosh(xi))
    vz = V0 * np.tanh(xi)
    om_a = kz + alpha * Az1
    om_f = kz - alpha * Az1
    i = np.arange(N)
    ip = (i + 1) % N
    im = (i - 1 + N) % N
    live = np.ones(N, bool) if xi_cut is None else (np.abs(xi) <= xi_cut)
    ib, iex, iez, ia, iqa, iqb = (k * N for k in range(6))
    rows, cols, vals = [], [], []

    def add(r, c, v, mask):
        v = np.broadcast_to(np.asarray(v, dtype=complex), (N,))
        rows.append(r[mask])
        cols.append(c[mask])
        vals.append(v[mask])

    L = live
    add(ib + i, iex + i, -1j * om_f, L)                 # Faraday: -i Om_F ex
    add(ib + i, iez + i, np.full(N, -1.0 / dx), L)      #   + (ez[i+1] - ez[i]) / dx
    add(ib + i, iez + ip, np.full(N, 1.0 / dx), L & live[ip])
    add(iex + i, ib + i, -1j * om_a, L)                 # Ampere x: -i Om_A b
    add(iez + i, ib + i, np.full(N, 1.0 / dx), L)       # Ampere z: (b[i] - b[i-1]) / dx
    add(iez + i, ib + im, np.full(N, -1.0 / dx), L & live[im])
    add(iez + i, iqa + i, vz, L)                        #   + vz (qA - qB)
    add(iez + i, iqb + i, -vz, L)
    add(ia + i, iez + i, np.full(N, -1.0), L)           # potential: -ez
    add(iqa + i, ia + i, 1j * alpha * vz, L)            # precession, beam A
    add(iqa + i, iqa + i, -1j * vz * om_a, L)
    add(iqb + i, ia + i, 1j * alpha * vz, L)            # precession, beam B
    add(iqb + i, iqb + i, 1j * vz * om_a, L)
    for blk in range(6):                                # cut rows: zero (eigenvalue 0)
        add(blk * N + i, blk * N + i, np.zeros(N), ~L)
    if xi_sponge is not None:
        excess = np.abs(xi) - xi_sponge
        damp = excess > 0
        for blk in range(6):
            add(blk * N + i, blk * N + i, -sigma_sponge * excess + 0j, damp)
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = np.concatenate(vals)
    return sp.coo_matrix((v, (r, c)), shape=(6 * N, 6 * N), dtype=complex).tocsr()


def solve(kz, alpha, V0, EPS, NX, Lx=LX_DEFAULT, n_eigs=10, sigma=None, xi_sponge=None, sigma_sponge=5.0,
          xi_cut=None):
    """(eigenvalues by decreasing real part, eigenvectors as columns), or
    (None, None) when the Arnoldi iteration returns nothing."""
    L = build_operator(kz, alpha, V0, EPS, NX, Lx, xi_sponge, sigma_sponge, xi_cut)
    if sigma is None:
        gw = wkb_growth_rate(kz, alpha, V0)
        sigma = gw * 0.55 if (np.isfinite(gw) and gw > 0) else 0.1
    try:
        vals, vecs = spla.eigs(L, k=n_eigs, sigma=sigma, which="LM", maxiter=10000, tol=1e-10)
    except spla.ArpackNoConvergence as e:
        vals, vecs = e.eigenvalues, e.eigenvectors
        if len(vals) == 0:
            return None, None
    order = np.argsort(-vals.real)
    return vals[order], vecs[:, order]


def profiles(vec, NX):
    """(b, ex, ez, a, qA, qB), each of length NX."""
    return tuple(vec[k * NX:(k + 1) * NX] for k in range(6))


def is_localised(vec, NX, EPS, Lx=LX_DEFAULT, frac=0.15, xi_inner=None):
    """|b| peaks at |xi| < 1.15 xi_inner (default 5) and is below frac x peak
    at both ends of the domain."""
    if xi_inner is None:
        xi_inner = 5.0
    amp = np.abs(vec[0:NX])
    ip = int(np.argmax(amp))
    xi_peak = abs(ip * Lx / NX - Lx / 2.0) / EPS
    return bool(xi_peak < 1.15 * xi_inner and max(amp[0], amp[NX - 1]) < frac * amp[ip])


def growing_mode(vals, vecs, NX, EPS, Lx=LX_DEFAULT, xi_sponge=None, require_localised=False):
    """Index of the selected growing eigenvalue: the first (largest real part)
    with Re > 1e-5 that is localised inside the sponge (xi_inner = xi_sponge,
    else 5); without one, the first with Re > 1e-5 unless require_localised.
    None when nothing qualifies."""
    if vals is None:
        return None
    xi_inner = xi_sponge if xi_sponge is not None else 5.0
    for j, v in enumerate(vals):
        if v.real > 1e-5 and is_localised(vecs[:, j], NX, EPS, Lx, xi_inner=xi_inner):
            return j
    if require_localised:
        return None
    for j, v in enumerate(vals):
        if v.real > 1e-5:
            return j
    return None


def growth_rate(kz, alpha, V0, EPS, NX, Lx=LX_DEFAULT, xi_sponge=None, sigma_sponge=5.0, xi_cut=None,
                require_localised=False):
    """(gamma complex, eigenvector) of the selected growing mode, or (nan, None)."""
    vals, vecs = solve(kz, alpha, V0, EPS, NX, Lx, xi_sponge=xi_sponge, sigma_sponge=sigma_sponge, xi_cut=xi_cut)
    j = growing_mode(vals, vecs, NX, EPS, Lx, xi_sponge, require_localised)
    if j is None:
        return complex("nan"), None
    return vals[j], vecs[:, j]


def seed_profiles(vec, NX):
    """(6, NX) float32: By2, Ex2, Ez2, Az2, Q2A, Q2B, phase-aligned so that
    Az peaks real and positive, divided by max |Az|."""
    b, ex, ez, a, qa, qb = profiles(vec, NX)
    ip = int(np.argmax(np.abs(a)))
    phase = np.exp(-1j * np.angle(a[ip]))
    scale = np.max(np.abs(a))
    return np.stack([(f * phase).real / scale for f in (b, ex, ez, a, qa, qb)]).astype(np.float32)


def seed_filename(kz_int, alpha, V0, EPS, xi_sponge, Lz=LZ_DEFAULT, Lx=LX_DEFAULT):
    lz_tag = f"_lz{Lz:.4f}" if abs(Lz - LZ_DEFAULT) > 1e-6 else ""
    lx_tag = f"_lx{Lx:.4f}" if abs(Lx - LX_DEFAULT) > 1e-6 else ""
    eps_tag = f"_EPS{EPS:.3f}" if abs(EPS - EPS_DEFAULT) > 1e-6 else ""
    return f"eigenmode_seed_kz{kz_int}_a{alpha:.2f}_V{V0:.3f}{eps_tag}_sp{xi_sponge:.1f}{lz_tag}{lx_tag}.bin"


def write_seed(path, prof):
    """int32 {6, NX} header, then the (6, NX) float32 profiles."""
    prof = np.asarray(prof, dtype=np.float32)
    with open(path, "wb") as f:
        np.array([prof.shape[0], prof.shape[1]], dtype=np.int32).tofile(f)
        prof.tofile(f)


def read_seed(path):
    with open(path, "rb") as f:
        hdr = np.fromfile(f, dtype=np.int32, count=2)
        prof = np.fromfile(f, dtype=np.float32, count=int(hdr[0]) * int(hdr[1]))
    return prof.reshape(int(hdr[0]), int(hdr[1]))


def export_seed(directory, kz_int, alpha, V0, EPS, NX, xi_sponge, Lz=LZ_DEFAULT, Lx=LX_DEFAULT,
                sigma_sponge=5.0, xi_cut=None):
    """Solve at kz = kz_int 2 pi / Lz, write the eigen-seed file of the
    selected growing mode into `directory`; returns (path, gamma)."""
    import os

    kz = kz_int * 2.0 * np.pi / Lz
    sp_arg = xi_sponge if xi_sponge and xi_sponge > 0 else None
    gamma, vec = growth_rate(kz, alpha, V0, EPS, NX, Lx, sp_arg, sigma_sponge, xi_cut)
    if vec is None:
        raise RuntimeError(f"no growing mode at kz={kz}, alpha={alpha}, V0={V0}")
    path = os.path.join(directory, seed_filename(kz_int, alpha, V0, EPS, xi_sponge, Lz, Lx))
    write_seed(path, seed_profiles(vec, NX))
    return path, gamma


def main(argv=None):
    """Command line of ym_eigenmode.py (:263-444): growth rates per kz, and
    with --export-seed the seed files, --save-profiles the profile .npz."""
    import argparse
    import os

    ap = argparse.ArgumentParser(description="SU(2) KH linear eigenmodes (log-cosh background)")
    ap.add_argument("--alpha", type=float, default=2.0)
    ap.add_argument("--V0", type=float, default=0.1)
    ap.add_argument("--EPS", type=float, default=EPS_DEFAULT)
    ap.add_argument("--kz", type=int, nargs="+", default=[1, 2, 3, 4, 5, 6])
    ap.add_argument("--NX", type=int, default=384)
    ap.add_argument("--Lz", type=float, default=LZ_DEFAULT)
    ap.add_argument("--Lx", type=float, default=LX_DEFAULT)
    ap.add_argument("--sigma", type=float, default=None)
    ap.add_argument("--xi-sponge", type=float, default=10.0)
    ap.add_argument("--sigma-sponge", type=float, default=5.0)
    ap.add_argument("--xi-cut", type=float, default=None)
    ap.add_argument("--export-seed", action="store_true")
    ap.add_argument("--save-profiles", action="store_true")
    ap.add_argument("--out", default=".")
    a = ap.parse_args(argv)
    xi_sp = a.xi_sponge if a.xi_sponge > 0 else None
    print(f"{'kz':>4} {'gamma':>10} {'Im':>10} {'gamma_WKB':>10}")
    for k in a.kz:
        kz = k * 2.0 * np.pi / a.Lz
        gw = wkb_growth_rate(kz, a.alpha, a.V0)
        vals, vecs = solve(kz, a.alpha, a.V0, a.EPS, a.NX, a.Lx, sigma=a.sigma, xi_sponge=xi_sp,
                           sigma_sponge=a.sigma_sponge, xi_cut=a.xi_cut)
        j = growing_mode(vals, vecs, a.NX, a.EPS, a.Lx, xi_sp)
        if j is None:
            print(f"{k:>4}  no growing mode")
            continue
        g = vals[j]
        print(f"{k:>4} {g.real:>10.4f} {g.imag:>10.4f} {gw:>10.4f}")
        if a.save_profiles:
            b, ex, ez, az, qa, qb = profiles(vecs[:, j], a.NX)
            np.savez(os.path.join(a.out, f"eigenmode_kz{k}_a{a.alpha:.2f}_V{a.V0:.2f}_EPS{a.EPS:.2f}.npz"),
                     x=grid(a.NX, a.EPS, a.Lx)[0], b=b, ex=ex, ez=ez, a=az, qA=qa, qB=qb, gamma=g, gamma_wkb=gw,
                     alpha=a.alpha, V0=a.V0, EPS=a.EPS, kz_int=k, kz_phys=kz, Lz=a.Lz)
        if a.export_seed:
            path = os.path.join(a.out, seed_filename(k, a.alpha, a.V0, a.EPS, a.xi_sponge, a.Lz, a.Lx))
            write_seed(path, seed_profiles(vecs[:, j], a.NX))
            print(f"       seed -> {path}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
