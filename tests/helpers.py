"""Conversions between the device tensors and the oracle's host families (tests only)."""
import numpy as np

from oracle.blocks import Fam, MpoSite


def fam_from_device(t):
    lay = t.layout
    return Fam(lay.row.as_dict(), lay.col.as_dict(), lay.shifts, t.to_blocks())


def device_from_fam(f, layout=None):
    from bondupdategalerkin_amd.blocks import BlockTensor, Layout, Leg

    if layout is None:
        layout = Layout(Leg.make(f.row.items()), Leg.make(f.col.items()), f.shifts)
    return BlockTensor.from_blocks(layout, f.blocks)


def oracle_mpo_site(site):
    """Oracle MpoSite from a product MPOSite (same entries)."""
    return MpoSite(site.cl, site.cr, zip(site.a, site.b, site.sout, site.sin, site.coeff))


def rel_err(a: Fam, b: Fam) -> float:
    num = 0.0
    keys = set(a.blocks) | set(b.blocks)
    for k in keys:
        x = a.blocks.get(k)
        y = b.blocks.get(k)
        if x is None:
            x = np.zeros_like(y)
        if y is None:
            y = np.zeros_like(x)
        num += np.sum(np.abs(x - y) ** 2)
    den = max(a.norm() ** 2, b.norm() ** 2, 1e-300)
    return float(np.sqrt(num / den))


def random_legs(rng, L, dmax, symmetric=True):
    """Random U(1) link legs {charge: dim} for a chain of L spin-1/2 sites, total charge 0."""
    legs = []
    for b in range(L + 1):
        if not symmetric:
            legs.append({0: int(min(2 ** min(b, L - b), dmax))})
            continue
        qs = [q for q in range(-b, b + 1, 2) if abs(q) <= L - b]
        if b in (0, L):
            legs.append({0: 1})
            continue
        leg = {q: int(rng.integers(1, dmax + 1)) for q in qs}
        legs.append(leg)
    return legs
