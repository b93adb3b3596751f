"""CPU restatement of the router's segment MLP (1 -> 16 -> 1, ReLU) — TEST INFRASTRUCTURE.

Used only by tests/ to check the device forward/backward kernels (eirene_amd/csrc/
stm_learned.hip: mlp_accumulate, k_mlp_grad).  Parameter layout: w1[0:16], b1[16:32],
w2[32:48], b2[48].  loss = mean_i 0.5 * (y_i - t_i)^2.  Pinned in tests by central finite
differences in float64 and by the closed form of a linear model.
"""
import numpy as np

H = 16


def forward(p, x):
    p = np.asarray(p, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    z = np.outer(x, p[:H]) + p[H:2 * H]
    h = np.maximum(z, 0.0)
    return h @ p[2 * H:3 * H] + p[3 * H], z, h


def loss_and_grad(p, x, t):
    p = np.asarray(p, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    n = max(len(x), 1)
    y, z, h = forward(p, x)
    e = y - t
    loss = 0.5 * np.sum(e * e) / n
    dy = e / n
    g = np.zeros(3 * H + 1)
    g[3 * H] = dy.sum()
    g[2 * H:3 * H] = h.T @ dy
    dz = (z > 0) * np.outer(dy, p[2 * H:3 * H])
    g[:H] = dz.T @ x
    g[H:2 * H] = dz.sum(axis=0)
    return loss, g
