"""QNM amplitudes of a ringdown: per-time inversion of the spherical
multipoles h_lm(t) (and their time derivatives) into spheroidal quasi-normal
mode amplitudes, then a sliding-window search for the stretch where those
amplitudes are most stable.

Follows the procedure of QNM_Extraction/Scripts/Extract_Ringdown_MoreModes.py
(preparesystem :13-79, getlinearsystem :81-84, main :96-209),
QNM_Extraction/Library/ringdownlib.py (loadtrajectory :5-46, loadwaveforms
:48-117, mulmlpnp :119-135, loadqnmtable :137-149, getqnmfrequencies
:151-188) and QNM_Extraction/Scripts/Postprocess.py:19-164, on numpy arrays:

* light-ring time t_lr: the trajectory row whose radius is closest to the
  prograde equatorial light ring 2(1 + cos(2/3 arccos(-|a|)));
* h_lm on a 0.1 M grid (linear interpolation), from t_lr - 50 M on, and
  dh_lm/dt by second-order central differences;
* at each time t, solve  sum_{l',+-} mu_{l l' +-} e^{w_{l'+-} (t - t_lr)} C_{l'+-} = h_l
  together with the same rows multiplied by w for dh_l/dt, where mu are the
  spherical-spheroidal overlaps and w = -i omega_lmn - 1/tau of the Kerr QNMs
  (tables interpolated linearly in a);
* amplitude |C| and unwrapped phase, averaged over the 50 M window (starting
  in the first 200 M) with the smallest summed relative standard deviation.

The QNM frequency and overlap tables are the reference's data files
(Tables/kerr/n{n}l{l}m{m}.dat, Tables/swsh/l{l}m{m}lp{l'}np{n}.dat, with "mm"
for negative m); QnmTables reads them from a directory or takes the arrays
directly.
"""
import math
import os

import numpy as np


def light_ring_radius(a):
    """Prograde equatorial light ring of a Kerr hole of spin a (M = 1)."""
    return 2.0 * (1.0 + math.cos((2.0 / 3.0) * math.acos(-abs(a))))


def light_ring_time(traj_t, traj_r, a):
    """Time of the trajectory row closest to the light ring (ringdownlib.py:38)."""
    return float(traj_t[int(np.argmin(np.abs(np.asarray(traj_r) - light_ring_radius(a))))])


def orbital_peak_time(traj_t, phidot):
    """Time of the largest orbital frequency on the trajectory's uniform grid
    (ringdownlib.py:15-29)."""
    t = np.asarray(traj_t)
    grid = np.arange(t[0], t[-1], t[1] - t[0])
    return float(grid[int(np.argmax(np.interp(grid, t, phidot)))])


def _mlabel(m):
    return f"mm{abs(m)}" if m < 0 else f"m{m}"


class QnmTables:
    """Kerr QNM frequencies and spherical-spheroidal overlaps, each a table
    in the spin a, interpolated linearly.  `arrays` maps a table's file name
    (e.g. "kerr/n1l2m2.dat", "swsh/l2m2lp3np0.dat") to its rows."""

    def __init__(self, root=None, arrays=None):
        self.root = root
        self.arrays = dict(arrays or {})

    def _table(self, name):
        if name not in self.arrays:
            if self.root is None:
                raise KeyError(f"table {name} not available")
            path = os.path.join(self.root, name)
            if not os.path.isfile(path):
                raise FileNotFoundError(path)
            self.arrays[name] = np.atleast_2d(np.loadtxt(path))
        return self.arrays[name]

    def qnm(self, l, m, a, n=0):
        """Complex QNM frequency omega_r + i omega_i (omega_i < 0) of the
        (l, m, n) mode; negative m maps to the opposite spin
        (ringdownlib.py:137-149)."""
        while m < 0:
            a, m = -a, -m
        name = f"kerr/n{n + 1}l{l}m{m}.dat" if a >= 0 else f"kerr/n{n + 1}l{l}mm{m}.dat"
        d = self._table(name)
        x = abs(a)
        return complex(np.interp(x, d[:, 0], d[:, 1]), np.interp(x, d[:, 0], d[:, 2]))

    def overlap(self, m, l, lp, a, n=0):
        """mu_{m l l' n}(a) = <Y_lm, S_l'mn> (ringdownlib.py:119-135)."""
        name = f"swsh/l{l}m{m}lp{lp}np{n}.dat" if m >= 0 else f"swsh/l{l}mm{abs(m)}lp{lp}np{n}.dat"
        d = self._table(name)
        if d.shape[0] == 1:
            return complex(d[0, 1], d[0, 2])
        return complex(np.interp(a, d[:, 0], d[:, 1]), np.interp(a, d[:, 0], d[:, 2]))


def mode_list(m, mmax):
    """Spherical l's fitted for azimuthal m: max(2,|m|) .. |m| + mmax."""
    return list(range(max(2, abs(m)), abs(m) + mmax + 1))


def spheroidal_modes(m, mmax):
    """(l', m') of the fitted QNMs, the column order of the amplitudes: for each
    l', co- and counter-rotating (for m = 0 the counter-rotating one carries
    l' < 0) (Extract_Ringdown_MoreModes.py:29-34)."""
    am = abs(m)
    cols = [(l, s) for l in mode_list(m, mmax) for s in (am, -am)]
    if m == 0:
        cols = [(l * (-1) ** i, 0) for i, (l, _) in enumerate(cols)]
    return cols


def system(m, a, mmax, tables):
    """Overlap matrix MU and exponent matrix W of the linear system
    (Extract_Ringdown_MoreModes.py:13-79): rows are h_l (l ascending) then
    dh_l/dt, columns the spheroidal modes of spheroidal_modes()."""
    ls = mode_list(m, mmax)
    rows = [(l, d) for d in (0, 1) for l in ls]
    cols = spheroidal_modes(m, mmax)
    n = len(rows)
    MU = np.zeros((n, n), dtype=complex)
    W = np.zeros((n, n), dtype=complex)
    for i, (lv, d) in enumerate(rows):
        for j, (lp, mp) in enumerate(cols):
            mu_ = tables.overlap(mp, lv, abs(lp), a)
            co = (mp * m > 0) or (m == 0 and lp > 0)
            mu = mu_.conjugate() if co else (-1) ** lv * mu_
            q = tables.qnm(abs(lp), mp, a)
            om_r = -q.real if (m == 0 and lp < 0) else q.real
            om_i = abs(q.imag)
            if mp * m < 0:
                w = 1j * om_r - om_i
            else:
                w = -1j * om_r - om_i
            if d:
                mu = mu * w
            MU[i, j] = mu
            W[i, j] = w
    return MU, W


def resample(t, cols, step=None):
    """Linear interpolation of data columns onto np.arange(t0, t_last, step)
    (the last sample excluded); step defaults to the data's first spacing
    (ringdownlib.py:93-103)."""
    t = np.asarray(t)
    step = t[1] - t[0] if step is None else step
    grid = np.arange(t[0], t[-1], step)
    return grid, [np.interp(grid, t, c) for c in cols]


def spherical_multipoles(wave, m, mmax):
    """h_l = h+ - i hx for the fitted l's from waveform rows in the reference's
    hm*.dat layout (t, then (h+, hx) for l = |m| .. |m|+6), on the file's own
    uniform grid and then on a 0.1 M grid (Extract_Ringdown_MoreModes.py:
    103-139)."""
    wave = np.asarray(wave)
    am = abs(m)
    cols = []
    for l in mode_list(m, mmax):
        ip = 1 + 2 * (l - am)
        cols += [wave[:, ip], wave[:, ip + 1]]
    t1, c1 = resample(wave[:, 0], cols)
    t2, c2 = resample(t1, c1, 0.1)
    h = np.array([c2[2 * i] - 1j * c2[2 * i + 1] for i in range(len(c2) // 2)])
    return t2, h


def fit_amplitudes(wave, m, a, mmax, t_lr, tables):
    """Spheroidal amplitudes C(t) from t_lr - 50 M to the end of the data:
    returns (times, C[times, 2 * n_l]) in the column order of
    spheroidal_modes()."""
    t, h = spherical_multipoles(wave, m, mmax)
    dt = t[1] - t[0]
    i0 = int(np.searchsorted(t, t_lr - 50.0, side="left"))
    hs = h[:, i0:-1]
    dh = np.array([np.gradient(x, dt) for x in hs])
    rhs = np.concatenate([hs, dh], axis=0).T                     # [T, 2L]
    ts = t[i0:-1]
    MU, W = system(m, a, mmax, tables)
    A = MU[None, :, :] * np.exp(W[None, :, :] * (ts - t_lr)[:, None, None])
    C = np.linalg.solve(A, rhs[:, :, None])[:, :, 0]
    return ts, C


def stable_window(times, C, window=50.0, span=250.0):
    """Mean amplitude and phase (mod 2 pi) over the `window`-long stretch,
    starting within the first span - window of the record, whose summed
    relative standard deviation of the amplitudes is smallest
    (Postprocess.py:52-96).  Returns (amp, phase, (t_low, t_high))."""
    times = np.asarray(times)
    amp = np.abs(C).T
    ang = np.unwrap(np.angle(C), axis=0).T
    best = (math.inf, None, None, None)
    i = 0
    t_high = 0.0
    stop = min(times[-1], times[0] + span)
    while t_high < stop:
        t_low = times[i]
        t_high = t_low + window
        hits = np.flatnonzero(np.abs(times - t_high) < 1e-10)
        if hits.size != 1:
            raise ValueError(f"no sample at t = {t_high}")
        j = int(hits[0])
        mean = amp[:, i:j].mean(axis=1)
        score = float(np.sum(amp[:, i:j].std(axis=1) / mean))
        if score < best[0]:
            best = (score, mean, ang[:, i:j].mean(axis=1) % (2 * np.pi), (t_low, t_high))
        i += 1
    return best[1], best[2], best[3]


def write_spheroidal(path, times, C, m, mmax):
    """The reference's spheroidal_coefficients_*.out layout
    (Extract_Ringdown_MoreModes.py:180-205)."""
    ls = mode_list(m, mmax)
    with open(path, "w") as f:
        f.write("#[0] time\n")
        for i, l in enumerate(ls):
            f.write(f"#[{4 * i + 1}] realamp_l{l}{_mlabel(m)}\n#[{4 * i + 2}] imagamp_l{l}{_mlabel(m)}\n")
            f.write(f"#[{4 * i + 3}] realamp_l{l}{_mlabel(-m)}\n#[{4 * i + 4}] imagamp_l{l}{_mlabel(-m)}\n")
        for t, row in zip(times, C):
            f.write("%.2f" % t + "".join("\t%.5e\t%.5e" % (z.real, z.imag) for z in row) + "\n")


def read_spheroidal(path):
    """(times, C) from a spheroidal_coefficients_*.out file."""
    d = np.loadtxt(path)
    return d[:, 0], d[:, 1::2] + 1j * d[:, 2::2]


def write_excitation(path, m, mmax, rows):
    """The reference's excitation_coefficients_*.dat layout
    (Postprocess.py:112-160); rows = [(theta_fin, sign, amp[], phase[]), ...]."""
    ls = mode_list(m, mmax)
    pos, neg = ("", "prime") if m >= 0 else ("prime", "")
    with open(path, "w") as f:
        f.write("#[0] thetafin\n#[1] sign_thetafindot\n")
        for i, l in enumerate(ls):
            f.write(f"#[{4 * i + 2}] l{l}{_mlabel(abs(m))}{pos}_amp\n#[{4 * i + 3}] l{l}{_mlabel(abs(m))}{pos}_phase\n")
            f.write(f"#[{4 * i + 4}] l{l}{_mlabel(-abs(m))}{neg}_amp\n#[{4 * i + 5}] l{l}{_mlabel(-abs(m))}{neg}_phase\n")
        for thf, sign, amp, ph in rows:
            f.write("%.2f\t%d" % (thf, sign) + "".join("\t%.4e\t%.4e" % (x, y) for x, y in zip(amp, ph)) + "\n")


def main(argv=None):
    """Command-line form of Extract_Ringdown_MoreModes.py followed by
    Postprocess.py for one trajectory: writes spheroidal_coefficients_*.out
    and excitation_coefficients_*.dat with the reference's names."""
    import argparse
    ap = argparse.ArgumentParser(description="QNM amplitudes of hm*.dat ringdown waveforms")
    ap.add_argument("--tables", required=True, help="folder holding kerr/ and swsh/ tables")
    ap.add_argument("--trajfilepath", required=True)
    ap.add_argument("--wavefilepaths", nargs="+", required=True)
    ap.add_argument("--modeindex", type=int, nargs="+", required=True)
    ap.add_argument("--mmax", type=int, required=True)
    ap.add_argument("--spinvalue", type=float, required=True)
    ap.add_argument("--thinc", type=int, required=True)
    ap.add_argument("--thf", required=True, help="e.g. 090.0_n (final polar angle and sign of its rate)")
    ap.add_argument("--outputdir", default=".")
    args = ap.parse_args(argv)
    tables = QnmTables(root=args.tables)
    traj = np.loadtxt(args.trajfilepath)
    t_lr = light_ring_time(traj[:, 0], traj[:, 1], args.spinvalue)
    files = {}
    for p in args.wavefilepaths:
        b = os.path.basename(p)
        files[-int(b[3]) if b[3] != "_" else int(b[2])] = p
    a_tag = ("%.2f" % args.spinvalue)[-2:]
    thf = float(args.thf.split("_")[0])
    sign = -1 if args.thf.split("_")[-1].startswith("n") else 1
    for m in args.modeindex:
        ts, C = fit_amplitudes(np.loadtxt(files[m]), m, args.spinvalue, args.mmax, t_lr, tables)
        sph = os.path.join(args.outputdir, f"spheroidal_coefficients_{_mlabel(m)}_a{a_tag}_thinc"
                                           f"{args.thinc:03d}_thf{args.thf}.out")
        write_spheroidal(sph, ts, C, m, args.mmax)
        tf, Cf = read_spheroidal(sph)
        amp, ph, _ = stable_window(tf, Cf)
        rows = [(thf, sign, amp, ph)]
        if args.thinc in (0, 180):
            rows = [(thf, -1, amp, ph), (thf, 1, amp, ph)]
        write_excitation(os.path.join(args.outputdir, f"excitation_coefficients_a{a_tag}_thinc{args.thinc:03d}_h"
                                                      f"{_mlabel(m)}.dat"), m, args.mmax, rows)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
