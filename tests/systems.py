"""Seeded test systems shared by the CPU and GPU tests."""
import numpy as np

from tis_amd import topology as T


def hairpin64(seed=0, coords="helix"):
    seq, pairs = T.config0_chain()
    s = T.build_rna_system([(seq, pairs)], name="c0")
    rng = np.random.default_rng(seed)
    x = T.helix_coords(s, rng) if coords == "helix" else T.compact_coords(s, rng, box=36, min_dist=2.6)
    return s, x


def compact_allpairs(seed=1, n=40, ions=False, periodic=False, cutoff_ele=99.0):
    rng = np.random.default_rng(seed)
    seq = T.random_sequence(rng, n)
    s = T.build_rna_system([(seq, [])], all_pair_hbonds=True,
                           ions=(["Mg"] * 3 + ["K"] * 5 + ["Cl"] * 4) if ions else None,
                           box=(60.0, 60.0, 60.0) if periodic else None, cutoff_ele=cutoff_ele,
                           name=f"compact{seed}")
    x = T.compact_coords(s, rng, box=28, min_dist=2.4)
    return s, x


def hairpin_design(seed, n):
    rng = np.random.default_rng(seed)
    stem = int(rng.integers(4, 11))
    loop = int(rng.integers(3, 9))
    seq, pairs = T.hairpin_design(rng, n, stem, loop)
    s = T.build_rna_system([(seq, pairs)], name=f"hp{seed}")
    x = T.helix_coords(s, rng)
    return s, x


def parity_cases():
    out = [("hairpin64_helix",) + hairpin64(0, "helix"),
           ("hairpin64_compact",) + hairpin64(3, "compact")]
    for k, kw in enumerate([{}, {"cutoff_ele": 2.0}, {"ions": True}, {"ions": True, "periodic": True}]):
        out.append((f"compact{k}",) + compact_allpairs(seed=10 + k, **kw))
    return out


def force_error(f, fo):
    """Per-bead relative force error |dF_i| / max(|F_i|, floor), floor = 1e-3 max|F|."""
    nf = np.linalg.norm(fo, axis=1)
    floor = 1e-3 * max(nf.max(), 1e-12)
    return np.linalg.norm(f - fo, axis=1) / np.maximum(nf, floor)
