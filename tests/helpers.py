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


def free_fermion_sz(L, t, occupied, J=1.0):
    """<S^z_j>(t) of the XX chain from the occupied sites' single-particle propagator
    (xx_free_fermion_sz, test/common/free_fermion.jl:37-40)."""
    import scipy.linalg as sla

    h = np.diag(np.full(L - 1, J / 2), 1) + np.diag(np.full(L - 1, J / 2), -1)
    U = sla.expm(-1j * t * h)
    return np.array([sum(abs(U[j, k]) ** 2 for k in occupied) - 0.5 for j in range(L)])


def mps_distance(a, b):
    """||a - b|| / ||b|| of two U(1) MPS given as lists of families, without densifying: the
    difference is one MPS whose bonds are the direct sums of the two, swept left to right by
    sector-wise QR, so no overlaps cancel (one family per site, blocks keyed (slice, row charge))."""
    def sweep(sites_a, sites_b, sign):
        carry = None                     # {charge: matrix} on the direct-sum bond
        L = len(sites_a)
        for i in range(L):
            fa = sites_a[i]
            fb = sites_b[i] if sites_b is not None else None
            rows = {}
            shifts = fa.shifts
            for k, sh in enumerate(shifts):
                for q in set(fa.row) | (set(fb.row) if fb is not None else set()):
                    qa = fa.blocks.get((k, q))
                    qb = fb.blocks.get((k, q)) if fb is not None else None
                    if qa is None and qb is None:
                        continue
                    ra, ca = (fa.row.get(q, 0), fa.col.get(q + sh, 0))
                    rb, cb = ((fb.row.get(q, 0), fb.col.get(q + sh, 0)) if fb is not None
                              else (0, 0))
                    first, last = i == 0, i == L - 1
                    R = (max(ra, rb) if first else ra + rb)
                    C = (max(ca, cb) if last else ca + cb)
                    blk = np.zeros((R, C), dtype=complex)
                    if qa is not None:
                        blk[:ra, :ca] += qa
                    if qb is not None:
                        r0 = 0 if first else ra
                        c0 = 0 if last else ca
                        blk[r0:r0 + rb, c0:c0 + cb] += (sign if first else 1.0) * qb
                    if carry is not None:
                        if q not in carry:
                            continue
                        blk = carry[q] @ blk
                    rows.setdefault(q + sh, []).append(blk)
            carry = {}
            for qc, parts in rows.items():
                M = np.vstack(parts)
                carry[qc] = np.linalg.qr(M, mode="r")
        return float(np.sqrt(sum(np.linalg.norm(m) ** 2 for m in carry.values())))

    return sweep(a, b, -1.0) / _norm_mps(b)


def _norm_mps(fams):
    """||psi|| of a U(1) MPS by a sector-wise QR sweep."""
    carry = None
    for f in fams:
        rows = {}
        for (k, q), blk in f.blocks.items():
            if carry is not None:
                if q not in carry:
                    continue
                blk = carry[q] @ blk
            rows.setdefault(q + f.shifts[k], []).append(blk)
        carry = {qc: np.linalg.qr(np.vstack(p), mode="r") for qc, p in rows.items()}
    return float(np.sqrt(sum(np.linalg.norm(m) ** 2 for m in carry.values())))
