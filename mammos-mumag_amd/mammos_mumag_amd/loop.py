"""Hysteresis-loop driver on the GPU path (scripts/loop.py:71-142 solve / loop).

Per field value: one ``Model.minimize`` (the NCG of minimizers.py:139-190 with
``total_eg`` and ``M_inv``), then ``M.h``; a snapshot is written whenever M.h has
dropped by more than ``mstep`` since the last one (the first step always saves), and
the record ``[counter, hext, M.h, E]`` is kept.  The potential ``u`` and the full
gradient (the reference's ``alt_args``) carry over between field steps; the
preconditioner uses the initial magnetisation throughout (loop.py:40-55 with 153-158).
"""
import time

import numpy as np

from . import _lib
from .fileio import write_dat, write_vtu

FMAX = np.finfo(np.float64).max


def run_loop(model, params, m0, name=None, save=True, max_steps=None, on_step=None,
             save_callback=None):
    """Returns (rec, stats_dict, m_final_host)."""
    hstart, hfinal, hstep = params["hstart"], params["hfinal"], params["hstep"]
    if max_steps is None:
        max_steps = int(abs(hfinal - hstart) / abs(hstep)) + 1
    mfinal, mstep, hdir = params["mfinal"], params["mstep"], params["h"]
    m = model.to_device(m0)
    m_ref = m.clone()
    u = model.empty1()
    gF = model.empty3()
    stats = _lib.Stats()
    mh = FMAX
    last_saved = FMAX
    counter = 0
    hext = hstart
    rec = []
    idx = 0
    while mh > mfinal and idx < max_steps:
        t0 = time.perf_counter()
        energy = model.minimize(m, hext, hdir, params, m_ref, u, gF, stats)
        mh = model.compute_mh(m, hdir)
        if (last_saved - mh) > mstep:
            counter += 1
            last_saved = mh
            if save_callback is not None:
                save_callback(counter, m, u)
            elif save and name is not None:
                write_vtu(f"{name}_{counter:04d}.vtu", model.coords, model.tets, model.elem_tags,
                          {"m": m.cpu().numpy().reshape(-1, 3), "u": u.cpu().numpy()})
        rec.append([counter, hext, mh, energy])
        if on_step is not None:
            on_step(idx, hext, mh, energy, stats.as_dict(), time.perf_counter() - t0)
        hext = hext + hstep
        idx += 1
    if name is not None:
        write_dat(f"{name}.dat", rec)
    return rec, stats.as_dict(), m.cpu().numpy()
