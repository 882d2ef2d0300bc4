"""Stand-alone decks: the reference's case directories driving the HIP solver.

A case directory holds ``datin/input.*`` (read by readinput,
src/readwrite.F90:466-716) and ``datin/controller`` (readcont,
src/readwrite.F90:730-760).  ``run_deck`` builds the solver from them, sets
the flow type's initial field, advances ``maxstep + 1`` RK3 steps as
steploop does (src/mainloop.F90:37-214: the loop runs while nstep <=
maxstep), and writes ``flowstate.dat`` as statout does
(src/statistic.F90:776-857): a header line (listinit,
src/utility.F90:310-362) and, for nstep = 1..maxstep, one line of the
statistics rkfirst takes at stage 1 of that step (listwrite,
src/utility.F90:374-434).

    python -m astr_gpu_amd.deck CASEDIR [--input datin/input.tgv] [--maxstep N] [--out DIR]

Built flow types: ``tgv`` (tgvini), ``hit`` (the seeded turbulence
generator of this package, same spectrum as the reference's) and
``riem2d`` (riem2dini with the example's user boundary condition).  The
deck must be nondimensional with a generated grid (lreadgrid f), no
species, no turbulence model and no immersed boundary; the other cases
raise ``DeckError``.

Checkpoints: every feqchkpt steps (at the start of a step nstep with
mod(nstep, feqchkpt) = 0, after the first, as rkfirst calls writechkpt,
src/mainloop.F90:668-673) the state goes to outdat/flowfield.npz and
outdat/auxiliary.txt (astr_gpu_amd.checkpoint; npz in place of the
reference's HDF5).  A deck with lrestart t reads them back
(readcheckpoint, src/initialisation.F90:46-50) and continues from that
nstep; flowstate.dat is resumed as listinit does (src/utility.F90:310-362):
the lines before that nstep are kept and the run writes on from there.
"""
from __future__ import annotations

import argparse
import math
import os
import re
import sys


class DeckError(ValueError):
    pass


# ------------------------------------------------------------------ reading

def _value_lines(path):
    """The value records of a deck: readinput and readcont skip the
    comment lines between them with '(/)' formats; every comment line of
    the reference's decks starts with '#'."""
    with open(path) as f:
        for line in f:
            s = line.strip()
            if s and not s.startswith("#"):
                yield s


def _items(line):
    """List-directed items of one record: separated by commas and/or
    blanks."""
    return [t for t in re.split(r"[,\s]+", line.strip()) if t]


def _real(t):
    # Fortran reals may carry a d exponent (0.49d0, 1.d-4)
    return float(t.lower().replace("d", "e"))


def _logical(t):
    # list-directed logicals: t, .true., T... (the first letter after an
    # optional period decides)
    s = t.lower().lstrip(".")
    if s[:1] == "t":
        return True
    if s[:1] == "f":
        return False
    raise DeckError(f"not a logical: {t}")


def read_input(path):
    """readinput (src/readwrite.F90:466-716), the build without COMB: the
    records in order, extra items on a record ignored as a list-directed
    read ignores them."""
    recs = list(_value_lines(path))
    it = iter(recs)

    def nxt():
        try:
            return _items(next(it))
        except StopIteration:
            raise DeckError(f"{path}: too few records") from None

    d = {}
    d["flowtype"] = nxt()[0]
    r = nxt()
    d["ia"], d["ja"], d["ka"] = (int(v) for v in r[:3])
    r = nxt()
    d["lihomo"], d["ljhomo"], d["lkhomo"] = (_logical(v) for v in r[:3])
    r = nxt()
    keys = ["nondimen", "diffterm", "lfilter", "lreadgrid", "lfftk", "limmbou", "ltimrpt"]
    for k, v in zip(keys, r):
        d[k] = _logical(v)
    d["lrestart"] = _logical(nxt()[0])
    r = nxt()
    d["alfa_filter"], d["kcutoff"] = _real(r[0]), int(r[1])
    r = nxt()
    if d["nondimen"]:
        d["ref_tem"], d["reynolds"], d["mach"] = (_real(v) for v in r[:3])
    else:
        d["ref_tem"], d["ref_vel"], d["ref_len"], d["ref_den"] = (_real(v) for v in r[:4])
    r = nxt()
    d["conschm"], d["difschm"], d["rkscheme"] = r[0], r[1], r[2]
    r = nxt()
    d["recon_schem"], d["lchardecomp"] = int(r[0]), _logical(r[1])
    d["bfacmpld"], d["shkcrt"] = _real(r[2]), _real(r[3])
    r = nxt()
    d["num_species"] = int(r[0])
    r = nxt()
    d["turbmode"], d["iomode"] = r[0], r[1] if len(r) > 1 else ""
    # src/readwrite.F90:561-575: bctype per face, 41 followed by twall,
    # 411 by xslip and twall, 421 by xslip
    bct, twall = [], [1.0] * 6
    for n in range(6):
        r = nxt()
        bct.append(int(r[0]))
        if bct[-1] == 41 and len(r) > 1:
            twall[n] = _real(r[1])
        elif bct[-1] == 411 and len(r) > 2:
            d["xslip"], twall[n] = _real(r[1]), _real(r[2])
        elif bct[-1] == 421 and len(r) > 1:
            d["xslip"] = _real(r[1])
    d["bctype"] = bct
    d["twall"] = twall
    d["ninit"] = int(nxt()[0])
    r = nxt()
    d["spg"] = [int(v) for v in r[:6]]
    d["gridfile"] = nxt()[0]
    return d


def read_controller(path):
    """readcont (src/readwrite.F90:730-760): output switches, step counts
    and frequencies, the time step."""
    it = iter(list(_value_lines(path)))
    try:
        r = _items(next(it))
        c = dict(zip(["lwsequ", "lwslic", "lavg", "lcracon"], (_logical(v) for v in r[:4])))
        r = _items(next(it))
        keys = ["maxstep", "feqchkpt", "feqwsequ", "feqslice", "feqlist", "feqavg"]
        c.update(zip(keys, (int(v) for v in r[:6])))
        c["deltat"] = _real(_items(next(it))[0])
    except StopIteration:
        raise DeckError(f"{path}: too few records") from None
    return c


def find_input(casedir, name=None):
    """The input file of a case: name (relative to the case directory, as
    the reference takes it on its command line), else datin/input.dat
    (readinput's default), else the single datin/input.* there."""
    if name:
        return os.path.join(casedir, name)
    datin = os.path.join(casedir, "datin")
    p = os.path.join(datin, "input.dat")
    if os.path.exists(p):
        return p
    cands = sorted(f for f in os.listdir(datin) if f.startswith("input."))
    if len(cands) != 1:
        raise DeckError(f"{datin}: expected one input.* file, found {cands}")
    return os.path.join(datin, cands[0])


# ------------------------------------------------------------------ mapping

CONVECTION = {"643c": 0, "543e": 1, "543c": 2}
GRADIENT = {"643c": 0, "643e": 1}
FLOWTYPES = ("tgv", "hit", "riem2d")
BOX_LENGTHS = {"tgv": (2 * math.pi,) * 3, "hit": (2 * math.pi,) * 3, "riem2d": (1.0, 1.0, 1.0)}
# the conditions of src/bc.F90 the library applies (gcnscbc, zeroextrap,
# symmetry, the noslip walls, outflow, fixbc)
BUILT_BCTYPES = {23, 50, 60, 41, 42, 21, 31, 421, 411}


def params_from_deck(d):
    """AstrParams of a deck.  A homogeneous axis is periodic (npdc 3); any
    other axis lies on one domain bounded at both ends (npdc 4); a
    two-dimensional box (ka = 0) keeps npdck 3."""
    import astr_gpu_amd as ag

    def need(cond, msg):
        if not cond:
            raise DeckError(msg)

    need(d["flowtype"] in FLOWTYPES, f"flow type {d['flowtype']!r} is not built ({FLOWTYPES})")
    need(d["nondimen"], "dimensional decks are not built")
    need(not d["lreadgrid"], "grids read from a file are not built (lreadgrid f)")
    need(not d["limmbou"], "immersed boundaries are not built")
    need(d["num_species"] == 0, "species are not built")
    need(d["turbmode"] == "none", "turbulence models are not built")
    need(d["rkscheme"] == "rk3", "only rk3 is built")
    need(d["ninit"] == 0, "only the flow type's own initial field (ninit 0) is built")
    need(d["conschm"] in CONVECTION, f"convection scheme {d['conschm']!r} is not built")
    need(d["difschm"] in GRADIENT, f"gradient scheme {d['difschm']!r} is not built")
    dims = (d["ia"], d["ja"], d["ka"])
    two_d = dims[2] == 0
    homo = (d["lihomo"], d["ljhomo"], d["lkhomo"])
    npdc = [3 if h else 4 for h in homo]
    if two_d:
        npdc[2] = 3
    for n, b in enumerate(d["bctype"]):
        if b == 1 or (b == 0 and d["flowtype"] == "riem2d"):
            continue
        need(b in BUILT_BCTYPES,
             f"bctype {b} on face {n + 1} is not built ({sorted(BUILT_BCTYPES)}, 1, and 0 "
             f"for riem2d)")
    return ag.default_params(dims, npdci=npdc[0], npdcj=npdc[1], npdck=npdc[2],
                             convection=CONVECTION[d["conschm"]],
                             deriv_explicit=GRADIENT[d["difschm"]],
                             recon_schem=d["recon_schem"], lchardecomp=int(d["lchardecomp"]),
                             lfilter=int(d["lfilter"]), diffterm=int(d["diffterm"]),
                             alfa_filter=d["alfa_filter"], reynolds=d["reynolds"],
                             mach=d["mach"], ref_tem=d["ref_tem"], bfacmpld=d["bfacmpld"],
                             shkcrt=d["shkcrt"])


# ------------------------------------------------------------------ writing

def fortran_e(x, w=20, d=13, e=2):
    """Fortran's Ew.dEe edit descriptor for output: sign, '0.', d digits
    rounded to nearest, 'E', the exponent sign and e digits, right-justified
    in w columns (asterisks when it does not fit)."""
    if not math.isfinite(x):
        s = "NaN" if math.isnan(x) else ("-Infinity" if x < 0 else "Infinity")
        return s.rjust(w)[-w:]
    if x == 0.0:
        mant, exp = "0" * d, 0
    else:
        m, ex = f"{abs(x):.{d - 1}e}".split("e")
        mant, exp = m.replace(".", ""), int(ex) + 1
    sexp = f"{'-' if exp < 0 else '+'}{abs(exp):0{e}d}"
    if len(sexp) > e + 1:
        return "*" * w
    s = f"{'-' if x < 0 else ''}0.{mant}E{sexp}"
    if len(s) > w:
        s = s.replace("0.", ".", 1)
    return s.rjust(w) if len(s) <= w else "*" * w


def flowstate_header(names):
    """listinit's first line: each name in (1X,A20), right-justified."""
    return "".join(" " + n.rjust(20)[-20:] for n in names)


def flowstate_line(nstep, time, values):
    """listwrite's line: (1X,I20,n(1X,E20.13E2)) of nstep, time and the
    values."""
    return " " + str(int(nstep)).rjust(20) + "".join(" " + fortran_e(v) for v in (time, *values))


# ------------------------------------------------------------------ running

def initial_state(d, p, seed=1):
    import astr_gpu_amd as ag

    ft = d["flowtype"]
    if ft == "tgv":
        return ag.tgv_state(p, lengths=BOX_LENGTHS[ft])
    if ft == "riem2d":
        return ag.riemann2d_state(p)
    return ag.hit_state(p, seed)


def resumed_lines(path, nstep):
    """listinit with lrestart (src/utility.F90:310-362): the header and the
    lines of flowstate.dat before the first one whose nstep is >= nstep
    (None if the file does not exist: a new file with its header)."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        lines = f.readlines()
    kept = lines[:1]
    for ln in lines[1:]:
        it = _items(ln)
        if not it or int(it[0]) >= nstep:
            break
        kept.append(ln)
    return kept


def run_deck(casedir, out=None, maxstep=None, seed=1, device=0, log=sys.stdout, input_name=None):
    """Run a case directory; returns the path of the flowstate.dat written
    (in out, default the case directory)."""
    import astr_gpu_amd as ag
    from astr_gpu_amd.checkpoint import read_checkpoint, write_checkpoint

    d = read_input(find_input(casedir, input_name))
    c = read_controller(os.path.join(casedir, "datin", "controller"))
    if maxstep is not None:
        c["maxstep"] = int(maxstep)
    p = params_from_deck(d)
    ft = d["flowtype"]
    s = ag.AstrSolver(p, device=device)
    outdat = os.path.join(out or casedir, "outdat")
    try:
        s.upload_geom(*ag.box_metrics(p, lengths=BOX_LENGTHS[ft]))
        if any(b != 1 for b in d["bctype"]):
            lx, ly, lz = BOX_LENGTHS[ft]
            s.set_lengths(lx, ly, lz if p.km > 0 else 0.0)
            s.set_twall(d["twall"])
            if "xslip" in d:
                s.set_xslip(d["xslip"])
            s.set_bc(tuple(d["bctype"]), ag.UDF_RIEMANN2D if ft == "riem2d" else 0)
        nstep0 = 0
        if d["lrestart"]:
            nstep0 = read_checkpoint(s, outdat)
        else:
            st = initial_state(d, p, seed)
            s.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
            s.time = 0.0
        names = ["nstep", "time"] + (["kenergy", "enstophy", "dissipation"]
                                     if ft in ("tgv", "hit") else
                                     ["maxq1", "maxq2", "maxq3", "maxq4", "maxq5"])
        path = os.path.join(out or casedir, "flowstate.dat")
        dt, feqlist = c["deltat"], max(1, c["feqlist"])
        kept = resumed_lines(path, nstep0) if d["lrestart"] else None
        with open(path, "w") as fs:
            if kept is None:
                fs.write(flowstate_header(names) + "\n")
            else:
                fs.writelines(kept)
            for nstep in range(nstep0, c["maxstep"] + 1):
                if nstep != nstep0 and c["feqchkpt"] > 0 and nstep % c["feqchkpt"] == 0:
                    write_checkpoint(s, outdat, nstep)
                t = s.time
                if ft in ("tgv", "hit"):
                    stats = s.time_integration_rk(dt, stats=True)
                    qmx = None
                else:
                    # rkfirst's maxmincal sees q after stage 1's boucon and
                    # qswap, which the step repeats unchanged
                    s.boucon()
                    s.qswap()
                    qmx = s.qmax()
                    stats = None
                    s.time_integration_rk(dt)
                if nstep == nstep0:
                    continue  # rkfirst writes nothing on the first call
                vals = stats if qmx is None else qmx
                fs.write(flowstate_line(nstep, t, vals) + "\n")
                if log is not None and nstep % feqlist == 0:
                    log.write(f"  {nstep:7d} {t:13.6e} " + " ".join(f"{v:13.6e}" for v in vals) + "\n")
        return path
    finally:
        s.close()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("casedir")
    ap.add_argument("--input", default=None,
                    help="input file relative to casedir (default datin/input.dat, or the only "
                         "datin/input.*)")
    ap.add_argument("--out", default=None, help="directory of flowstate.dat (default: casedir)")
    ap.add_argument("--maxstep", type=int, default=None, help="override the controller's maxstep")
    ap.add_argument("--seed", type=int, default=1, help="hit: seed of the turbulence generator")
    args = ap.parse_args(argv)
    path = run_deck(args.casedir, out=args.out, maxstep=args.maxstep, seed=args.seed,
                    input_name=args.input)
    print(f" << {path}")


if __name__ == "__main__":
    main()
