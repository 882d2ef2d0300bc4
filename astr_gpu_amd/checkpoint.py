"""Checkpoint and restart of a device state (writechkpt / readcheckpoint).

The reference writes a checkpoint from rkfirst every feqchkpt steps
(src/mainloop.F90:668-673): writechkpt -> writeflfed (src/readwrite.F90:
2376-2445, 1720-1918) puts nstep, time and the primitive fields ro, u1, u2,
u3, p, t on the nodes 0..im, 0..jm, 0..km into outdat/flowfield.h5, and
the namelist /restart/ (nstep, filenumb, fnumslic, ninflowslice, nsamples)
into outdat/auxiliary.txt.  A run with lrestart reads them back
(readcheckpoint, src/readwrite.F90:1378-1528, called from
src/initialisation.F90:46-50), stops if the two nstep values disagree, and
rebuilds q from the primitives (updateq).

The image has no HDF5 library, so the flow field goes into an npz archive
(outdat/flowfield.npz) holding the same variables under the same names,
Fortran-ordered arrays of the nodes 0..n.  auxiliary.txt is the reference's
namelist text.  Beside the reference's variables the archive holds the
device state as it is between two steps (q and the primitives with their
halos: the interior primitives are stale after the filter, so q rebuilt
from them is not the q the solver holds), under ``state_*``.  A restart
from an archive that has them resumes bit for bit; one without them (only
the reference's variables) takes the reference's path, q from the
primitives by fvar2q (src/fludyna.F90:449-533).

    write_checkpoint(solver, "outdat", nstep)
    nstep = read_checkpoint(solver, "outdat")
"""
from __future__ import annotations

import os
import re

import numpy as np

from .solver import HM

AUX_KEYS = ("nstep", "filenumb", "fnumslic", "ninflowslice", "nsamples")
STATE_KEYS = ("q", "rho", "vel", "prs", "tmp")


class CheckpointError(RuntimeError):
    pass


def write_auxiliary(path: str, nstep: int, filenumb: int = 0, fnumslic: int = 0,
                    ninflowslice: int = 0, nsamples: int = 0) -> None:
    """The namelist /restart/ as a Fortran list-directed namelist write."""
    vals = dict(nstep=nstep, filenumb=filenumb, fnumslic=fnumslic, ninflowslice=ninflowslice,
                nsamples=nsamples)
    with open(path, "w") as f:
        f.write("&RESTART\n")
        f.write(",\n".join(f" {k.upper()}={int(vals[k])}" for k in AUX_KEYS))
        f.write("\n /\n")


def read_auxiliary(path: str) -> dict:
    """read(16, nml=restart): name=value pairs between &restart and '/', any
    case, separated by commas or blanks; missing names keep 0."""
    with open(path) as f:
        text = f.read()
    m = re.search(r"&restart(.*?)/", text, re.IGNORECASE | re.DOTALL)
    if not m:
        raise CheckpointError(f"no &restart namelist in {path}")
    out = {k: 0 for k in AUX_KEYS}
    for name, val in re.findall(r"(\w+)\s*=\s*([-+]?\d+)", m.group(1)):
        if name.lower() in out:
            out[name.lower()] = int(val)
    return out


def _interior(a, n):
    return a[HM:HM + n[0] + 1, HM:HM + n[1] + 1, HM:HM + n[2] + 1]


def write_checkpoint(solver, folder: str, nstep: int, filenumb: int = 0) -> str:
    """Write folder/flowfield.npz and folder/auxiliary.txt from the solver's
    device state; returns the flow-field path."""
    os.makedirs(folder, exist_ok=True)
    p = solver.p
    n = (p.im, p.jm, p.km)
    st = solver.download_state()
    rec = {"nstep": np.int64(nstep), "time": np.float64(solver.time),
           "ro": _interior(st["rho"], n), "u1": _interior(st["vel"][..., 0], n),
           "u2": _interior(st["vel"][..., 1], n), "u3": _interior(st["vel"][..., 2], n),
           "p": _interior(st["prs"], n), "t": _interior(st["tmp"], n),
           "dims": np.array(n, dtype=np.int64)}
    for k in STATE_KEYS:
        rec["state_" + k] = st[k]
    path = os.path.join(folder, "flowfield.npz")
    tmp = path + ".tmp.npz"
    np.savez(tmp, **{k: np.asfortranarray(v) if isinstance(v, np.ndarray) else v
                     for k, v in rec.items()})
    os.replace(tmp, path)
    write_auxiliary(os.path.join(folder, "auxiliary.txt"), nstep, filenumb)
    return path


def read_checkpoint(solver, folder: str) -> int:
    """Restore the solver from folder/flowfield.npz and folder/auxiliary.txt;
    returns nstep.  Raises CheckpointError where readcheckpoint stops: the
    two nstep values disagree, or the box size differs."""
    aux = read_auxiliary(os.path.join(folder, "auxiliary.txt"))
    path = os.path.join(folder, "flowfield.npz")
    with np.load(path, allow_pickle=False) as z:
        nstep = int(z["nstep"])
        if nstep != aux["nstep"]:
            raise CheckpointError(f"flowfield.npz (nstep {nstep}) not consistent with "
                                  f"auxiliary.txt (nstep {aux['nstep']})")
        p = solver.p
        n = (p.im, p.jm, p.km)
        if tuple(int(v) for v in z["dims"]) != n:
            raise CheckpointError(f"checkpoint of a {tuple(z['dims'])} box, solver {n}")
        if all("state_" + k in z.files for k in STATE_KEYS):
            st = {k: np.asfortranarray(z["state_" + k]) for k in STATE_KEYS}
        else:
            st = state_from_primitives(p, z["ro"], z["u1"], z["u2"], z["u3"], z["p"], z["t"])
        solver.upload_state(st["q"], st["rho"], st["vel"], st["prs"], st["tmp"])
        solver.time = float(z["time"])
    return nstep


def state_from_primitives(p, ro, u1, u2, u3, prs, tmp) -> dict:
    """updateq's state from the primitives on the nodes: fvar2q with the
    temperature (src/fludyna.F90:449-533), halos zero (qswap fills them)."""
    from .solver import haloed_shape

    n = (p.im, p.jm, p.km)
    const1 = 1.0 / (p.gamma * (p.gamma - 1.0) * (p.mach * p.mach))
    s = haloed_shape(p)
    I = (slice(HM, HM + n[0] + 1), slice(HM, HM + n[1] + 1), slice(HM, HM + n[2] + 1))
    out = {"q": np.zeros(s + (5,), order="F"), "rho": np.zeros(s, order="F"),
           "vel": np.zeros(s + (3,), order="F"), "prs": np.zeros(s, order="F"),
           "tmp": np.zeros(s, order="F")}
    out["rho"][I] = ro
    for c, u in enumerate((u1, u2, u3)):
        out["vel"][I + (c,)] = u
    out["prs"][I] = prs
    out["tmp"][I] = tmp
    out["q"][I + (0,)] = ro
    out["q"][I + (1,)] = ro * u1
    out["q"][I + (2,)] = ro * u2
    out["q"][I + (3,)] = ro * u3
    out["q"][I + (4,)] = ro * (tmp * const1 + 0.5 * (u1 * u1 + u2 * u2 + u3 * u3))
    return out
