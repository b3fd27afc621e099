"""Batched many-molecule pipeline: grouped launches, autotuned tiles, HIP graph
capture and replay, checked per molecule against the single-molecule path
(itself checked against the oracle) and, for the smallest molecule, the
oracle directly."""
import numpy as np
import pytest

from oracle import ri_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12


def relf(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def D():
    from rest_tensors_amd import device
    assert device.device_count() >= 1, "no GPU visible"
    return device


def make_mols(D, nbs, seed0):
    mols = []
    for i, nb in enumerate(nbs):
        naux, nocc = 3 * nb, max(1, nb // 8)
        dA = D.DeviceArray(nb * nb * naux)
        D.gen_ri3ao(dA, nb, naux, seed=seed0 + 2 * i)
        Cm = O.gen_orthonormal(nb, seed=seed0 + 2 * i + 1)
        mols.append(dict(nbas=nb, naux=naux, nocc=nocc, ri3ao=dA,
                         C=D.DeviceArray.from_host(Cm),
                         D=D.DeviceArray(nb * nb), J=D.DeviceArray(nb * nb),
                         K=D.DeviceArray(nb * nb), B=D.DeviceArray(nocc * nb * naux), Cm=Cm))
    return mols


def single(D, m):
    nb, naux, nocc = m["nbas"], m["naux"], m["nocc"]
    dD, dJ, dK = (D.DeviceArray(nb * nb) for _ in range(3))
    dB = D.DeviceArray(nocc * nb * naux)
    D.density(m["C"], nb, nocc, dD)
    ej = D.rij(m["ri3ao"], nb, naux, dD, dJ)
    D.ao2mo_occ(m["C"], nb, nb, nocc, m["ri3ao"], naux, dB)
    ex = D.rik(dB, nocc, nb, naux, dK)
    return dict(D=dD.to_host((nb, nb)), J=dJ.to_host((nb, nb)), K=dK.to_host((nb, nb)),
                B=dB.to_host((nocc, nb, naux)), ej=ej, ex=ex)


def check_against(mols, ej, ex, refs):
    for i, (m, r) in enumerate(zip(mols, refs)):
        nb, naux, nocc = m["nbas"], m["naux"], m["nocc"]
        assert relf(m["D"].to_host((nb, nb)), r["D"]) < TOL, i
        assert relf(m["J"].to_host((nb, nb)), r["J"]) < TOL, i
        assert relf(m["B"].to_host((nocc, nb, naux)), r["B"]) < TOL, i
        K = m["K"].to_host((nb, nb))
        assert relf(K, r["K"]) < TOL, i
        assert np.array_equal(K, K.T)
        assert abs(ej[i] - r["ej"]) <= TOL * abs(r["ej"]), i
        assert abs(ex[i] - r["ex"]) <= TOL * abs(r["ex"]), i


def test_batch_eager_autotune_and_graph(D):
    from rest_tensors_amd.batch import Batch
    nbs = [64, 77, 130, 96, 201, 150]
    mols = make_mols(D, nbs, seed0=4000)
    refs = [single(D, m) for m in mols]
    b = Batch(mols)
    b.run()
    ej, ex = b.energies()
    check_against(mols, ej, ex, refs)
    # oracle on the smallest molecule
    m = mols[0]
    A = m["ri3ao"].to_host((64, 64, 192))
    _, J, e_j = O.rij(A, O.density(m["Cm"], 8))
    assert relf(m["J"].to_host((64, 64)), J) < TOL and abs(ej[0] - e_j) < TOL * abs(e_j)
    Kr, e_x, _ = O.rik(O.ao2mo_occ(m["Cm"], A, 8))
    assert relf(m["K"].to_host((64, 64)), Kr) < TOL and abs(ex[0] - e_x) < TOL * abs(e_x)

    b.autotune(reps=1)
    info = b.info()
    assert set(info) == {"density", "ao2mo_s1", "ao2mo_s2", "k_partials"}
    assert all(len(v["tile"]) == 2 and v["waves_m"] in (1, 2, 4) for v in info.values())
    assert info["k_partials"]["tile"][0] == info["k_partials"]["tile"][1]   # upper tiles
    b.run()
    ej, ex = b.energies()
    check_against(mols, ej, ex, refs)

    # graph: capture once, replay several times with eager work in between,
    # every replay checked (outputs are cleared before each replay)
    s = D.Stream()
    b.capture(s.handle)
    for rep in range(3):
        for mm in mols:
            mm["K"].zero()
            mm["J"].zero()
        D.synchronize()
        b.launch(s.handle)
        s.synchronize()
        ej, ex = b.energies(s.handle)
        check_against(mols, ej, ex, refs)
        single(D, mols[rep])   # unrelated eager work between replays
    s.destroy()


@pytest.mark.parametrize("s1,s2,k", [((128, 64, 2), (32, 128, 1), (64, 64, 2)),
                                     ((256, 32, 4), (64, 128, 2), (128, 128, 2)),
                                     ((128, 32, 4), (128, 128, 2), (64, 64, 2)),
                                     ((256, 64, 4), (64, 64, 2), (128, 128, 2)),
                                     ((256, 48, 4), (128, 128, 2), (64, 64, 2)),
                                     ((128, 48, 4), (64, 128, 2), (64, 64, 2))])
def test_batch_tn_tiles(D, s1, s2, k):
    """Every TN tile of stages 1, 2 and K, forced, against the single-molecule
    path; nbas and nocc not multiples of 16 (the padded internal layouts).
    nocc = 8, 16, 12, 31, 37: stage 1 (N = nocc) runs 1, 2 and 3 of a wave's
    16-column MFMA blocks, and whole waves past N (64-wide tiles) idle."""
    from rest_tensors_amd.batch import Batch
    mols = make_mols(D, [67, 130, 99, 250, 300], seed0=4200)
    refs = [single(D, m) for m in mols]
    b = Batch(mols)
    b.set_tiles("ao2mo_s1", *s1)
    b.set_tiles("ao2mo_s2", *s2)
    b.set_tiles("k_partials", *k)
    info = b.info()
    assert all(info[k]["engine"] == "tn" for k in ("ao2mo_s1", "ao2mo_s2", "k_partials"))
    b.run()
    ej, ex = b.energies()
    check_against(mols, ej, ex, refs)
    # new integrals: the padded copy follows after reload
    for i, m in enumerate(mols):
        D.gen_ri3ao(m["ri3ao"], m["nbas"], m["naux"], seed=4400 + i)
    refs2 = [single(D, m) for m in mols]
    b.reload()
    b.run()
    ej, ex = b.energies()
    check_against(mols, ej, ex, refs2)


@pytest.mark.parametrize("s1,s2", [((128, 48, 4), (48, 128, 1)), ((256, 32, 4), (32, 256, 1))])
def test_batch_generic_stages(D, monkeypatch, s1, s2):
    """Stages 1, 2 and K forced onto the generic grouped engine (narrow tiles,
    waves stacked along the wide side) against the single-molecule path."""
    from rest_tensors_amd.batch import Batch
    monkeypatch.setenv("RTA_BATCH_GROUPED", "1")
    mols = make_mols(D, [70, 131, 96], seed0=4300)
    refs = [single(D, m) for m in mols]
    b = Batch(mols)
    b.set_tiles("ao2mo_s1", *s1)
    b.set_tiles("ao2mo_s2", *s2)
    assert b.info()["ao2mo_s2"]["engine"] == "generic"
    b.run()
    ej, ex = b.energies()
    check_against(mols, ej, ex, refs)
