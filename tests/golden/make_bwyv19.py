"""Writes tests/golden/bwyv19_ref.npz: the reference's own Fortran routines
(oracle/_ref/libtisref.so, compiled from the reference's src/) evaluated on
the bwyv19 fixture (tests/golden/bwyv19/: ninfo/bwyv19.ninfo on
pdb/bwyv19_force_check.{rna,ion}.pdb), so that the oracle can be checked
against these vectors where the reference is absent.

Run from the repository root where oracle/_ref is built:
    python tests/golden/make_bwyv19.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.dirname(HERE)]

from oracle import RefLib  # noqa: E402
from systems import bwyv19  # noqa: E402

MT_SEED = 20261203
MT_WORDS = 4096


def pairs_digest(p):
    p = np.ascontiguousarray(np.asarray(p, dtype=np.int32)[np.lexsort(np.asarray(p).T[::-1])])
    return hashlib.sha256(p.tobytes()).hexdigest()


def main():
    s, x = bwyv19()
    R = RefLib()
    R.load(s)
    R.neighbor(x)
    out = {"xyz": x}
    exv = R.lists(0)
    ele, ele_coef = R.lists(1)
    out["exv_pairs"] = exv[np.lexsort(exv.T[::-1])]
    out["ele_count"] = np.int64(len(ele))
    out["ele_digest"] = np.array(pairs_digest(ele))
    out["ele_coef_sum"] = np.float64(np.sort(ele_coef).sum())
    out["hb_list"] = np.sort(R.lists(2))
    for t in ("bond", "angle", "stack", "exv", "dh"):
        out["f_" + t] = R.force_term(t)
    R.mt_seed(MT_SEED)
    out["f_hbond"] = R.force_term("hbond")
    st, ene = R.hb_state()
    out["hb_status"], out["hb_ene"] = st, ene
    out["mt_words"] = R.mt_raw(MT_SEED, MT_WORDS)
    v = np.random.default_rng(3).normal(size=x.shape)
    out["velo"] = v
    out["energy"] = R.energy(v)
    out["energy_unit0"] = R.energy_unit(v, 0)
    out["energy_unit1"] = R.energy_unit(v, 1)
    np.savez_compressed(os.path.join(HERE, "bwyv19_ref.npz"), **out)
    print("wrote bwyv19_ref.npz", {k: np.asarray(v).shape for k, v in out.items()})


if __name__ == "__main__":
    main()
