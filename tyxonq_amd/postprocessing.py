"""Result post-processing router for the simulator run() path.

Covers the methods of postprocessing/__init__.py:16-208 that touch this path:
no method ({"method": None, "result": None}) and the Pauli-Z aggregations
expval_pauli_term / expval_pauli_terms / expval_pauli_sum of
postprocessing/counts_expval.py.  Every term is read as a Z product on its
qubits, the letters being ignored, as there: basis rotations for X/Y letters
are part of the circuit (counts_expval.py:8-11, 67-76).

Four sources, in the reference's order of preference (counts_expval.py:
55-112):
  * counts (shots > 0): one parity column per term over the distinct sampled
    bitstrings, weighted by their counts;
  * shots = 0 with the run's GPU state at hand (Circuit.run passes it): all
    Z products in one fused FP64 reduction on the GPU (the Pauli-sum pass of
    libtqsv with x = 0 and one z mask per term), nothing copied to the host;
  * shots = 0 with host probabilities only: sign(parity) . p per term,
    vectorised over basis states instead of the reference's Python loop;
  * otherwise products of the single-qubit <Z_q> of ``expectations``.
Unknown methods return the reference's placeholder.
"""
from __future__ import annotations

from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np


def _counts_table(counts: Dict[str, int]) -> Tuple[np.ndarray, np.ndarray]:
    """(bits[m, n] uint8, weights[m] float) of a non-empty counts dict."""
    keys = list(counts)
    n = len(keys[0])
    bits = (np.frombuffer("".join(keys).encode("ascii"), dtype=np.uint8) - 48).reshape(-1, n)
    w = np.fromiter((counts[k] for k in keys), dtype=np.float64, count=len(keys))
    return bits, w


def _parity_mean(bits: np.ndarray, w: np.ndarray, total: float, idxs: Sequence[int]) -> float:
    par = np.zeros(bits.shape[0], dtype=np.uint8)
    for q in idxs:
        par ^= bits[:, int(q)]
    return float(np.dot(w, 1.0 - 2.0 * par)) / total


def term_expectation_from_counts(counts: Dict[str, int], idxs: Sequence[int]) -> float:
    """<Z^{idxs}> from big-endian bitstring counts (counts_expval.py:6-20):
    the count-weighted mean of (-1)^(parity of the bits at idxs)."""
    if not counts:
        return 0.0
    bits, w = _counts_table(counts)
    total = float(w.sum()) or 1.0
    return _parity_mean(bits, w, total, idxs)


def _split_item(entry: Any) -> Tuple[Tuple[Tuple[int, str], ...], Optional[float]]:
    """A Pauli-sum item as ((qubit, letter), ...) and its coefficient, if any.

    Accepted forms (counts_expval.py:23-32, plus a dict form): ``(term,
    coeff)`` with a numeric coeff, ``(term,)``, a bare term ``((q, P), ...)``,
    and ``{"pauli_terms"|"term": term, "coeff": c}``; (P, q) pairs in the dict
    form are flipped to (q, P)."""
    if isinstance(entry, dict):
        raw = entry.get("pauli_terms") or entry.get("term") or ()
        term = []
        for a, b in raw:
            term.append((int(a), str(b)) if isinstance(a, (int, np.integer)) else (int(b), str(a)))
        coeff = entry.get("coeff")
        return tuple(term), (None if coeff is None else float(coeff))
    head_is_term = isinstance(entry, tuple) and len(entry) > 0 and isinstance(entry[0], tuple)
    if not head_is_term:
        return tuple(entry), None
    coeff = entry[1] if len(entry) > 1 else None
    return tuple(entry[0]), (float(coeff) if isinstance(coeff, (int, float)) else None)


def _z_mask(idxs: Sequence[int], n: int) -> int:
    m = 0
    for q in idxs:
        m ^= 1 << (n - 1 - int(q))
    return m


def _z_products_from_probs(probs: np.ndarray, masks: Sequence[int]) -> List[float]:
    """sum_i (-1)^{|i & mask|} p_i per mask, with the parity of i & mask taken
    by folding the masked index onto its low bit."""
    idx = np.arange(probs.size, dtype=np.uint64)
    out = []
    for m in masks:
        v = idx & np.uint64(m)
        for sh in (32, 16, 8, 4, 2, 1):
            v ^= v >> np.uint64(sh)
        out.append(float(np.dot(1.0 - 2.0 * (v & np.uint64(1)).astype(np.float64), probs)))
    return out


def _z_products_on_gpu(psi: Any, masks: Sequence[int], n: int) -> List[float]:
    from . import _native as nat

    if not masks:
        return []
    x = np.zeros(len(masks), dtype=np.uint64)
    z = np.asarray(masks, dtype=np.uint64)
    ham = nat.Ham(n, x, z, np.ones(len(masks)))
    try:
        _, _, per_term = ham.expval(psi, per_term=True)
    finally:
        ham.close()
    return [float(v) for v in per_term]


def expval_pauli_sum(counts: Optional[Dict[str, int]], items: Sequence[Any],
                     identity_const: float = 0.0, *, expectations: Optional[Dict[str, float]] = None,
                     probabilities: Optional[np.ndarray] = None,
                     num_qubits: Optional[int] = None, device_state: Any = None
                     ) -> Dict[str, Any]:
    """{"energy": identity_const + sum coeff * <Z-product>, "expvals": [...]}
    (counts_expval.py:87-112); the analytic branch is taken when there are
    expectations and no counts (counts_expval.py:55-84)."""
    parsed = [_split_item(e) for e in items]
    qsets = [[int(q) for q, _ in term] for term, _ in parsed]
    analytic = expectations is not None and not counts
    if analytic and num_qubits is not None and (device_state is not None
                                                or probabilities is not None):
        n = int(num_qubits)
        masks = [_z_mask(qs, n) for qs in qsets]
        if device_state is not None:
            expvals = _z_products_on_gpu(device_state, masks, n)
        else:
            expvals = _z_products_from_probs(
                np.asarray(probabilities, dtype=np.float64).reshape(-1), masks)
    elif analytic:
        expvals = [float(np.prod([float(expectations.get(f"Z{q}", 0.0)) for q in qs]))
                   for qs in qsets]
    elif counts:
        bits, w = _counts_table(counts)
        total = float(w.sum()) or 1.0
        expvals = [_parity_mean(bits, w, total, qs) for qs in qsets]
    else:
        expvals = [0.0 for _ in qsets]
    energy = float(identity_const) + sum(c * v for (_, c), v in zip(parsed, expvals)
                                         if c is not None)
    return {"energy": float(energy), "expvals": expvals}


def apply_postprocessing(result: Dict[str, Any], options: Optional[Dict[str, Any]] = None,
                         device_state: Any = None) -> Dict[str, Any]:
    """postprocessing/__init__.py:16-208 for the methods above.  device_state
    is the run's final GPU state, when the caller still holds it."""
    opts = dict(options or {})
    method = opts.get("method")
    post: Dict[str, Any] = {"method": method, "result": None}
    if not method:
        return post
    if method not in ("expval_pauli_term", "expval_pauli_terms", "expval_pauli_sum"):
        post["result"] = {"reason": f"unsupported method: {method}"}
        return post
    meta = result.get("result_meta", {}) or {}
    counts = result.get("result") or result.get("counts") or {}
    if method == "expval_pauli_term":
        idxs = opts.get("idxs") or opts.get("indices") or []
        post["result"] = term_expectation_from_counts(counts, idxs)
        return post
    if method == "expval_pauli_terms":
        terms = opts.get("terms") or meta.get("group_items") or []
        post["result"] = [term_expectation_from_counts(counts, [int(q) for q, _ in
                                                                _split_item(t)[0]])
                          for t in terms]
        return post
    expectations = result.get("expectations") or meta.get("expectations") or None
    probabilities = result.get("probabilities")
    if probabilities is None:
        probabilities = meta.get("probabilities")
    if probabilities is None and device_state is None:
        sv = result.get("statevector")
        sv = meta.get("statevector") if sv is None else sv
        if sv is not None:
            probabilities = np.abs(np.asarray(sv)) ** 2
    n = None
    for src in (result.get("metadata"), meta, meta.get("result_meta")):
        if isinstance(src, dict) and src.get("num_qubits") is not None:
            n = int(src["num_qubits"])
            break
    if n is None and probabilities is not None:
        n = int(np.asarray(probabilities).size).bit_length() - 1
    if n is None and device_state is not None:
        n = int(device_state.numel()).bit_length() - 1
    items = opts.get("items") or opts.get("group_items") or meta.get("group_items") or []
    ic = float(opts.get("identity_const", meta.get("identity_const", 0.0)))
    post["result"] = expval_pauli_sum(counts, items, ic, expectations=expectations,
                                      probabilities=probabilities, num_qubits=n,
                                      device_state=device_state)
    return post
