"""Float64 restatement of the learned proposal -- TEST INFRASTRUCTURE ONLY.

The learned proposal has no counterpart in the reference (it is this project's
addition to the device-resident evolution). This module states its features,
forward pass and cross-entropy gradients in numpy float64 so that the device
kernels can be checked against them; its own gradient is pinned by central
finite differences (tests/test_oracle.py). Only tests/ import it.
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT, RANKS, KEEP = 64, 64, 32, 30, 30
W1 = 0
B1 = W1 + IN * HID
W2 = B1 + HID
B2 = W2 + HID * HID
W3 = B2 + HID
B3 = W3 + HID * OUT
PARAMS = B3 + OUT


def bf16(a):
    """Round float32 values to bf16 (nearest even), returned as float32."""
    b = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)


def features(neighbors, labels, k=None):
    """Row features (64) and ranked voted labels (30, -1 past them)."""
    nbr = np.asarray(neighbors)
    lab = np.asarray(labels)
    rows = nbr.shape[0]
    k = nbr.shape[1] if k is None else k
    feat = np.zeros((rows, IN), np.float32)
    ranked = np.full((rows, RANKS), -1, np.int32)
    for i in range(rows):
        codes = [int(lab[j]) for j in nbr[i] if 0 <= j < rows]
        vals, cnt = np.unique(np.array(codes, np.int64), return_counts=True) if codes else (np.zeros(0, np.int64),
                                                                                            np.zeros(0, np.int64))
        order = sorted(range(len(vals)), key=lambda t: (-int(cnt[t]), int(vals[t])))
        own = int(lab[i])
        own_count = int(cnt[vals == own][0]) if np.any(vals == own) else 0
        own_rank = RANKS
        for j, t in enumerate(order[:RANKS]):
            feat[i, j] = np.float32(np.float32(cnt[t]) / np.float32(k))
            ranked[i, j] = vals[t]
            if vals[t] == own and own_rank == RANKS:
                own_rank = j
        top = int(cnt[order[0]]) if order else 0
        feat[i, RANKS + own_rank] = 1.0
        feat[i, 61] = np.float32(np.float32(top - own_count) / np.float32(k))
        feat[i, 62] = 1.0
    return feat, ranked


def unpack(params, round_weights=True):
    p = np.asarray(params, np.float32)
    r = bf16 if round_weights else (lambda a: a)
    return dict(w1=r(p[W1:B1]).astype(np.float64).reshape(IN, HID), b1=p[B1:W2].astype(np.float64),
                w2=r(p[W2:B2]).astype(np.float64).reshape(HID, HID), b2=p[B2:W3].astype(np.float64),
                w3=r(p[W3:B3]).astype(np.float64).reshape(HID, OUT), b3=p[B3:PARAMS].astype(np.float64))


def valid_outputs(feat):
    f = np.asarray(feat)
    v = np.zeros((f.shape[0], OUT), bool)
    v[:, :RANKS] = (f[:, :RANKS] > 0) & (f[:, RANKS:2 * RANKS] == 0)
    v[:, KEEP] = True
    return v


def _forward(w, x, valid):
    h1p = x @ w["w1"] + w["b1"]
    h1 = np.maximum(h1p, 0.0)
    h2p = h1 @ w["w2"] + w["b2"]
    h2 = np.maximum(h2p, 0.0)
    z = h2 @ w["w3"] + w["b3"]
    zm = np.where(valid, z, -np.inf)
    e = np.where(valid, np.exp(zm - zm.max(axis=1, keepdims=True)), 0.0)
    prob = e / e.sum(axis=1, keepdims=True)
    return h1, h2, prob


def forward(params, feat):
    """Probabilities [rows x 32] in float64 (masked outputs 0)."""
    x = bf16(feat).astype(np.float64)
    return _forward(unpack(params), x, valid_outputs(feat))[2]


def loss_and_grads_of(w, feat, targets):
    """Mean cross-entropy and its gradient with respect to the (already
    rounded) weights and the biases, as a flat vector in the parameter layout."""
    x = bf16(feat).astype(np.float64)
    t = np.asarray(targets)
    n = x.shape[0]
    h1, h2, prob = _forward(w, x, valid_outputs(feat))
    loss = -np.mean(np.log(prob[np.arange(n), t]))
    dz = prob.copy()
    dz[np.arange(n), t] -= 1.0
    dz /= n
    dh2 = (dz @ w["w3"].T) * (h2 > 0)
    dh1 = (dh2 @ w["w2"].T) * (h1 > 0)
    g = np.zeros(PARAMS, np.float64)
    g[W1:B1] = (x.T @ dh1).ravel()
    g[B1:W2] = dh1.sum(axis=0)
    g[W2:B2] = (h1.T @ dh2).ravel()
    g[B2:W3] = dh2.sum(axis=0)
    g[W3:B3] = (h2.T @ dz).ravel()
    g[B3:PARAMS] = dz.sum(axis=0)
    return loss, g


def loss_and_grads(params, feat, targets):
    return loss_and_grads_of(unpack(params), feat, targets)
