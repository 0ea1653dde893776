"""The array operations of the numerics backends (numpy and pytorch), checked
against closed forms and identities rather than against each other alone:
exp of a Pauli rotation generator, factorisation round trips, derivatives of
known functions, sampling results that must be numpy arrays."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import tyxonq_amd as tq

X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
I2 = np.eye(2, dtype=np.complex128)


@pytest.fixture(params=["numpy", "pytorch"])
def K(request):
    return tq.get_backend(request.param)


def test_backend_objects_and_dtypes(K):
    # a named lookup makes a new instance; the configured one is shared
    assert tq.get_backend(K.name) is not K and tq.get_backend(K.name).name == K.name
    prev = tq.get_backend()
    try:
        assert tq.set_backend(K) is K and tq.get_backend() is K
        with tq.use_backend("numpy") as b:
            assert tq.get_backend() is b
        assert tq.get_backend() is K
    finally:
        tq.set_backend(prev)
    assert K.name in repr(K)
    assert K.set_dtype("complex64") == (K.complex64, K.float32) and K.rdtypestr == "float32"
    assert K.set_dtype("complex128") == (K.complex128, K.float64)
    with pytest.raises(ValueError):
        K.set_dtype("float16")
    if K.name == "numpy":
        assert K.complex128 is np.complex128 and K.int is np.int64
    else:
        assert K.complex64 is torch.complex64 and K.float64 is torch.float64


def test_rotation_generator_exponential(K):
    th = 0.731
    R = K.to_numpy(K.expm(K.asarray(-0.5j * th * X, dtype=np.complex128)))
    want = np.cos(th / 2) * I2 - 1j * np.sin(th / 2) * X
    np.testing.assert_allclose(R, want, atol=1e-12)


def test_kron_einsum_matmul_and_tensordot(K):
    ZX = K.kron(K.asarray(Z), K.asarray(X))
    assert K.to_numpy(ZX).shape == (4, 4)
    np.testing.assert_allclose(K.to_numpy(ZX)[2:, 2:], -X)
    v = K.asarray(np.arange(4, dtype=np.complex128))
    a = K.to_numpy(K.matmul(ZX, v))
    b = K.to_numpy(K.einsum("ij,j->i", ZX, v))
    np.testing.assert_allclose(a, b)
    np.testing.assert_allclose(K.to_numpy(K.dot(ZX, v)), a)
    A = K.asarray(np.arange(24.0).reshape(2, 3, 4))
    B = K.asarray(np.arange(12.0).reshape(4, 3))
    got = K.to_numpy(K.tensordot(A, B, axes=([1, 2], [1, 0])))
    np.testing.assert_allclose(got, np.tensordot(np.arange(24.0).reshape(2, 3, 4),
                                                 np.arange(12.0).reshape(4, 3), axes=([1, 2], [1, 0])))


def test_linear_algebra_round_trips(K):
    rng = np.random.default_rng(3)
    M = rng.normal(size=(5, 5)) + 1j * rng.normal(size=(5, 5))
    H = M + M.conj().T
    w, V = K.eigh(K.asarray(H))
    w, V = K.to_numpy(w), K.to_numpy(V)
    np.testing.assert_allclose(V @ np.diag(w) @ V.conj().T, H, atol=1e-10)
    U, s, Vh = (K.to_numpy(t) for t in K.svd(K.asarray(M)))
    np.testing.assert_allclose(U @ np.diag(s) @ Vh, M, atol=1e-10)
    b = rng.normal(size=5) + 0j
    x = K.to_numpy(K.solve(K.asarray(M), K.asarray(b)))
    np.testing.assert_allclose(M @ x, b, atol=1e-10)
    np.testing.assert_allclose(K.to_numpy(K.inv(K.asarray(M))) @ M, np.eye(5), atol=1e-10)
    ev = np.sort_complex(K.to_numpy(K.eig(K.asarray(H))[0]))
    np.testing.assert_allclose(ev.real, np.sort(w), atol=1e-10)
    assert float(K.to_numpy(K.norm(K.asarray(np.array([3.0, 4.0]))))) == pytest.approx(5.0)


def test_elementwise_and_shapes(K):
    a = K.asarray(np.array([[1.0, -4.0], [9.0, 16.0]]))
    np.testing.assert_allclose(K.to_numpy(K.sqrt(K.abs(a))), [[1, 2], [3, 4]])
    np.testing.assert_allclose(K.to_numpy(K.sign(a)), [[1, -1], [1, 1]])
    np.testing.assert_allclose(K.to_numpy(K.sum(K.square(a), axis=1)), [17, 337])
    assert float(K.to_numpy(K.mean(a))) == pytest.approx(5.5)
    np.testing.assert_allclose(K.to_numpy(K.log2(K.asarray(np.array([1.0, 8.0])))), [0, 3])
    t = K.reshape(K.arange(24), (2, 3, 4))
    assert K.to_numpy(K.moveaxis(t, 0, 2)).shape == (3, 4, 2)
    assert K.to_numpy(K.transpose(t)).shape == (4, 3, 2)
    assert K.to_numpy(K.transpose(t, (1, 0, 2))).shape == (3, 2, 4)
    s = K.stack([K.zeros((2,)), K.ones((2,))])
    np.testing.assert_allclose(K.to_numpy(s), [[0, 0], [1, 1]])
    np.testing.assert_allclose(K.to_numpy(K.concatenate([K.eye(2), K.eye(2)], axis=0)),
                               np.vstack([np.eye(2)] * 2))
    np.testing.assert_allclose(K.to_numpy(K.linspace(0, 1, 5)), np.linspace(0, 1, 5))
    c = K.asarray(np.array([1 + 2j, 3 - 4j]))
    np.testing.assert_allclose(K.to_numpy(K.real(K.conj(c))), [1, 3])
    np.testing.assert_allclose(K.to_numpy(K.imag(K.conj(c))), [-2, 4])
    np.testing.assert_allclose(K.to_numpy(K.outer(c, K.ones_like(c))), np.outer([1 + 2j, 3 - 4j], [1, 1]))
    np.testing.assert_allclose(K.to_numpy(K.diag(K.asarray(np.array([1.0, 2.0])))), np.diag([1.0, 2.0]))
    assert K.to_numpy(K.cast(K.asarray(np.array([1.7])), np.int64)).dtype == np.int64
    assert K.allclose(K.copy(c), c) and K.isclose(c, c).all()
    np.testing.assert_allclose(K.to_numpy(K.exp(K.asarray(np.array([0.0])))), [1.0])
    np.testing.assert_allclose(K.to_numpy(K.cos(K.zeros_like(K.asarray(np.array([2.0]))))), [1.0])
    np.testing.assert_allclose(K.to_numpy(K.sin(K.zeros((1,)))), [0.0])
    np.testing.assert_allclose(K.to_numpy(K.log(K.ones((1,)))), [0.0])


def test_sampling_helpers_return_numpy(K):
    r1, r2 = K.rng(11), K.rng(11)
    a, b = K.to_numpy(K.normal(r1, (3, 2))), K.to_numpy(K.normal(r2, (3, 2)))
    assert a.shape == (3, 2) and np.array_equal(a, b)
    idx = K.choice(K.rng(0), 4, size=200, p=np.array([0.0, 0.5, 0.5, 0.0]))
    assert isinstance(idx, np.ndarray) and set(np.unique(idx)) <= {1, 2}
    cnt = K.bincount(idx, minlength=4)
    assert isinstance(cnt, np.ndarray) and cnt.sum() == 200 and cnt[0] == cnt[3] == 0
    nz = K.nonzero(K.asarray(np.array([0.0, 2.0, 0.0, 1.0])))
    assert isinstance(nz, tuple) and isinstance(nz[0], np.ndarray)
    np.testing.assert_array_equal(nz[0], [1, 3])


def test_derivatives_of_known_functions(K):
    x0 = np.array([0.3, -1.2, 2.0])
    if K.name == "pytorch":
        f = lambda p: K.sum(K.square(K.sin(p)))
        jf = lambda p: K.sin(p)
    else:
        f = lambda p: np.sum(np.sin(p) ** 2)
        jf = np.sin
    val, g = K.value_and_grad(f)(x0)
    assert float(K.to_numpy(val)) == pytest.approx(np.sum(np.sin(x0) ** 2))
    np.testing.assert_allclose(K.to_numpy(g), np.sin(2 * x0), atol=1e-7)
    arg = K.asarray(x0) if K.name == "pytorch" else x0
    J = K.to_numpy(K.jacfwd(jf)(arg))
    np.testing.assert_allclose(J, np.diag(np.cos(x0)), atol=1e-6)
    sq = K.vmap(lambda v: K.sum(K.square(v)))
    rows = K.asarray(np.arange(6.0).reshape(3, 2))
    np.testing.assert_allclose(K.to_numpy(sq(rows)), [1, 13, 41])
    assert K.jit(f) is f


def test_torch_asarray_keeps_the_autograd_graph():
    K = tq.get_backend("pytorch")
    x = torch.tensor([0.5, 2.0], dtype=torch.float32, requires_grad=True)
    y = K.asarray(x, dtype=np.float64)
    assert y.dtype == torch.float64 and y.requires_grad
    K.sum(K.square(y)).backward()
    np.testing.assert_allclose(x.grad.numpy(), [1.0, 4.0])
    z = K.requires_grad(K.zeros((2,)))
    assert z.requires_grad and not K.detach(z).requires_grad
