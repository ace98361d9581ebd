"""Host-side model of the coarse-plane ring of the pair sweep with the
prolongation (the PRO variant of k_rbsweep3p in jsolver_amd/csrc/
jsv_kernels.hpp), in both march directions.

The sweep applies phi += P e to the black values entering its fine ring; e
comes from a ring of three coarse planes, slot I % 3 holding coarse plane I
(indexed by stored plane).  Stored fine plane f reads coarse planes
(f + 1) >> 1 .. (f + 2) >> 1.  An upward march adds the plane (f + 2) >> 1 of
the fine plane two steps ahead; a downward march (march plane u stored at
pl(u) = mz + 2 - u) adds (f + 1) >> 1.  The commit of a step is not separated
by a barrier from the previous step's corrections, which other threads may
still be computing, so the plane it replaces must be needed by neither.

The model restates the kernel's rules (ring set-up, per-step commit and
prefetch, when corrections are evaluated) and checks over many levels,
z-chunks and direction modes that every correction finds its coarse planes
resident, that no commit evicts a plane the previous or current step's
corrections read, that no commit is of a negative plane, and that each
commit is the plane prefetched one step before.  The GPU parity tests check
the kernel's bits against the oracle on fewer shapes.
"""
import itertools


def cmax(f):
    return (f + 2) >> 1


def cmin(f):
    return (f + 1) >> 1


def chunks(mz, zchunk, dn_in):
    """(z0, z1, dn) of every z-chunk as the kernel derives them from dn_in
    (bit 0: direction, bit 1: chunks of alternating direction whose bounds
    are the upward march's in stored planes)."""
    out = []
    n = (mz - 1 + zchunk - 1) // zchunk
    for c in range(n):
        alt = (dn_in & 2) != 0
        dn = (dn_in & 1) ^ ((c & 1) if alt else 0)
        z0 = 2 + c * zchunk
        z1 = min(z0 + zchunk, mz + 1)
        if alt and dn:
            z0, z1 = mz + 3 - z1, mz + 3 - z0
        out.append((z0, z1, dn))
    return out


def simulate(mz, z0, z1, dn):
    pl = (lambda u: mz + 2 - u) if dn else (lambda u: u)
    slot = {}                                   # ring slot -> coarse plane held

    def load(i):
        assert i >= 0, ("negative coarse plane", mz, z0, z1, dn, i)
        slot[i % 3] = i

    def needs(u):
        # corr_of evaluates only fine planes of unknowns (march plane u in 2..mz)
        if u < 2 or u > mz:
            return set()
        f = pl(u)
        return set(range(cmin(f), cmax(f) + 1))

    def resident(u):
        held = set(slot.values())
        missing = needs(u) - held
        assert not missing, ("coarse plane not in the ring", mz, z0, z1, dn, u, missing)

    if dn:
        chave = cmin(pl(z0 + 1))
        for i in range(chave, cmax(pl(z0 - 2)) + 1):
            load(i)
    else:
        chave = cmax(z0 + 1)
        for i in range((z0 - 1) >> 1, chave + 1):
            load(i)
    assert len(slot) == 3 or len(set(slot.values())) == len(slot)
    for u in (z0 - 2, z0 - 1, z0):             # the two ring planes loaded corrected, prep_corr(z0)
        resident(u)
    prefetched = None
    prev_corr = z0                              # fine plane of the last prep_corr

    def step(zs):
        nonlocal chave, prefetched, prev_corr
        want = cmin(pl(zs + 2)) if dn else cmax(zs + 2)
        if (want < chave) if dn else (want > chave):
            assert prefetched == want, ("commit of a plane not prefetched", mz, z0, z1, dn, zs)
            old = slot.get(want % 3)
            # the replaced plane: read by neither the previous step's
            # corrections nor this step's
            assert old is None or (old not in needs(prev_corr) and old not in needs(zs + 2)), \
                ("commit evicts a plane still read", mz, z0, z1, dn, zs, old)
            load(want)
            chave = want
        prefetched = cmin(pl(zs + 3)) if dn else cmax(zs + 3)
        resident(zs + 2)                        # prep_corr(zs + 2) at the end of the step
        prev_corr = zs + 2

    zs = z0 - 1
    while zs <= z1:
        step(zs)
        if zs + 1 <= z1:
            step(zs + 1)
        zs += 2


def test_coarse_ring_holds_every_correction_plane():
    n = 0
    for mz, zchunk in itertools.product([4, 6, 8, 14, 16, 30, 32, 62, 64, 126, 128, 256],
                                        [1, 2, 3, 4, 5, 7, 13, 20, 29, 64, 255]):
        if zchunk > mz - 1:
            continue
        for dn_in in (0, 1, 2, 3):
            for z0, z1, dn in chunks(mz, zchunk, dn_in):
                simulate(mz, z0, z1, dn)
                n += 1
    assert n > 1000


def test_model_catches_a_ring_that_moves_the_wrong_way():
    """The model is not vacuous: a downward march that kept the upward rule
    (adding cmax of the plane two steps ahead) fails it."""
    def wrong(mz, z0, z1):
        pl = lambda u: mz + 2 - u
        slot = {}
        chave = cmax(z0 + 1)
        for i in range((z0 - 1) >> 1, chave + 1):
            slot[i % 3] = i
        for zs in range(z0 - 1, z1 + 1):
            want = cmax(zs + 2)
            if want > chave:
                slot[want % 3] = want
                chave = want
            u = zs + 2
            if 2 <= u <= mz:
                f = pl(u)
                if not set(range(cmin(f), cmax(f) + 1)) <= set(slot.values()):
                    return False
        return True
    assert not all(wrong(64, z0, z1) for z0, z1, dn in chunks(64, 13, 3) if dn)
