"""The no-allocation leaf stage (knng_leaf_stage_device_noalloc) through the C ABI: the same
lists as knng_leaf_stage on two calls in a row (the second smaller, starting from existing lists),
and a clear error when its fixed scratch is too small."""
import numpy as np
import pytest

import oracle
import sigmod23ann_amd as K

pytestmark = pytest.mark.gpu


def _lowrank(n, d, seed, latent=24):
    rng = np.random.default_rng(seed)
    W = rng.normal(size=(latent, d))
    z = rng.normal(size=(n, latent)) / np.sqrt(1.0 + np.arange(latent))
    X = z @ W + 0.05 * rng.normal(size=(n, d))
    return (X / np.linalg.norm(X, axis=1, keepdims=True)).astype(np.float32)


def _run_noalloc(b, X, perm, ls, lz, state):
    n, d = X.shape
    k = b.prm.k
    dX = K.DeviceArray(X.nbytes)
    dX.upload(X)
    dP = K.DeviceArray(n * 4)
    dP.upload(perm)
    dI = K.DeviceArray(n * k * 4)
    dD = K.DeviceArray(n * k * 4)
    if state is not None:
        dI.upload(state[0])
        dD.upload(state[1])
    b.leaf_stage_device_noalloc(dX.ptr, n, d, d, dP.ptr, ls, lz, dI.ptr, dD.ptr, state is None)
    return dI.download((n, k), np.uint32), dD.download((n, k), np.float32)


def test_noalloc_leaf_stage_matches_allocating_stage_on_two_calls():
    k = 100
    kw = dict(k=k, leaf_max=400, small_max=3000, iterations=1, rounds=2, seed=17)
    o_prm = oracle.make_params(**kw)
    ref = K.Builder(K.params(**kw))
    b = K.Builder(K.params(**kw))
    # call 1: 60,000 points from empty lists (leaves up to 400, mixed bins)
    X1 = _lowrank(60_000, 100, 1)
    perm, ls, lz = oracle.split_round(X1, o_prm, 0)
    want = ref.leaf_stage(X1, perm, ls, lz)
    got = _run_noalloc(b, X1, perm, ls, lz, None)
    np.testing.assert_array_equal(got[1], want[1])
    np.testing.assert_array_equal(got[0], want[0])
    # call 2: fewer points, d = 64, starting from the lists of an earlier round, with leaves up
    # to 1500 points (the big-bucket kernel) in the tables
    kw2 = dict(kw, leaf_max=1500)
    o_prm2 = oracle.make_params(**kw2)
    ref.set_params(K.params(**kw2))
    b.set_params(K.params(**kw2))
    X2 = _lowrank(25_000, 64, 2)
    p0, l0, z0 = oracle.split_round(X2, o_prm2, 0)
    state = ref.leaf_stage(X2, p0, l0, z0)
    p1, l1, z1 = oracle.split_round(X2, o_prm2, 1)
    assert z1.max() > 512
    want = ref.leaf_stage(X2, p1, l1, z1, state=state)
    got = _run_noalloc(b, X2, p1, l1, z1, state)
    np.testing.assert_array_equal(got[1], want[1])
    np.testing.assert_array_equal(got[0], want[0])


def test_noalloc_leaf_stage_rejects_what_its_scratch_cannot_hold():
    b = K.Builder(K.params(k=10))
    n, d = 20_000_000, 4          # 80 MB of bounds alone: beyond the 64 MiB scratch
    dX = K.DeviceArray(n * d * 4)
    dP = K.DeviceArray(n * 4)
    dI = K.DeviceArray(n * 10 * 4)
    dD = K.DeviceArray(n * 10 * 4)
    with pytest.raises(K.KnngError) as e:
        b.leaf_stage_device_noalloc(dX.ptr, n, d, d, dP.ptr, [0], [2], dI.ptr, dD.ptr, True)
    assert e.value.code == K.ERR_UNSUPPORTED and "scratch" in str(e.value)
    with pytest.raises(K.KnngError) as e:     # rows that would need a padded copy
        b.leaf_stage_device_noalloc(dX.ptr, 1000, 6, 6, dP.ptr, [0], [2], dI.ptr, dD.ptr, True)
    assert e.value.code == K.ERR_UNSUPPORTED
