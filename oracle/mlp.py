"""FP64 oracle of the block-norm screener: features, the norm-product bound and the MLP.

Test infrastructure only: it restates, in numpy, what the library's screener kernels compute,
so that tests can check the device forward / backward passes and the bound.

The screener predicts log ||chi0s(I, J, R; tau)||_F for every output block before the block is
built.  LibRPA screens chi0 blocks only through the Cs threshold of its inputs; the learned
screener has no counterpart in the reference, so these functions cite the chi0 definition they
bound (LibRI labels, thirdparty/LibRI/include/RI/physics/RPA.hpp:104-138).

Features of one (block, tau) sample, in order (NFEAT = 8):
  0  |x_J + R A - x_I| / 10 bohr
  1  n_b(I) / 27,  2  n_b(J) / 27
  3  ln tau
  4  ln c_I,  5  ln c_J            c_A = sum over the neighbours n of A of ||C^A(n)||_F
  6  ln(||G+_IJ(R)||_F ||G-_IJ(R)||_F)
  7  ln(max ||G+||_F  max ||G-||_F)   (maxima over all blocks at this tau)

Norm-product bound (plain screener): every term of chi0s is a contraction of the form
<C^I_mu, G C^J_nu G'^T>, so ||chi0s(I,J,R)||_F <= c_I c_J (g+_IJ gm- + g-_IJ gm+ + 2 gm+ gm-),
with g the Frobenius norms of the IJ(R) blocks and gm the maxima over all blocks.

MLP: 8 -> 64 -> 64 -> 1 with SiLU, on standardised inputs and target:
  z1 = xs W1 + b1, h1 = silu(z1), z2 = h1 W2 + b2, h2 = silu(z2), y = h2 w3 + b3,
  loss = mean((y - ts)^2).  Parameter vector layout (NPARAM = 4801):
  W1[8][64], b1[64], W2[64][64], b2[64], w3[64], b3.
"""
import numpy as np

NFEAT = 8
H = 64
NPARAM = NFEAT * H + H + H * H + H + H + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + NFEAT * H].reshape(NFEAT, H); o += NFEAT * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    z1 = X @ W1 + b1
    h1 = silu(z1)
    z2 = h1 @ W2 + b2
    h2 = silu(z2)
    return h2 @ w3 + b3


def loss_and_grad(p, X, t):
    """Mean squared error over the samples and its gradient in the parameter layout."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    n = X.shape[0]
    z1 = X @ W1 + b1
    h1 = silu(z1)
    z2 = h1 @ W2 + b2
    h2 = silu(z2)
    y = h2 @ w3 + b3
    r = y - t
    loss = float(np.mean(r * r))
    dy = 2.0 * r / n
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    dz2 = np.outer(dy, w3) * dsilu(z2)
    gW2 = h1.T @ dz2
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2.T) * dsilu(z1)
    gW1 = X.T @ dz1
    gb1 = dz1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return loss, g


def init_params(seed=0):
    """Scaled-uniform initialisation (fan-in), zero biases."""
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (NFEAT, H)) / np.sqrt(NFEAT)
    W2 = rng.uniform(-1, 1, (H, H)) / np.sqrt(H)
    w3 = rng.uniform(-1, 1, H) / np.sqrt(H)
    return np.concatenate([W1.ravel(), np.zeros(H), W2.ravel(), np.zeros(H), w3, [0.0]])


def block_features(sys, tau, gp, gm, I, J, R, cnorm):
    """Feature vector of block (I, J, R) at tau; gp / gm: G+ / G- block norm dicts keyed like chi0s."""
    lat = np.asarray(sys.latvec)
    d = np.linalg.norm(sys.positions[J] + np.asarray(R, dtype=np.float64) @ lat - sys.positions[I])
    gpm = max(gp.values()) * max(gm.values())
    return np.array([d / 10.0, sys.nao[I] / 27.0, sys.nao[J] / 27.0, np.log(tau), np.log(cnorm[I]),
                     np.log(cnorm[J]), np.log(gp[(I, J, R)] * gm[(I, J, R)] + 1e-300), np.log(gpm + 1e-300)])


def cs_norm_sums(sys, cs):
    """c_A = sum over the folded Cs blocks C^A(K, R1) of their Frobenius norms."""
    c = np.zeros(sys.natom)
    for (I, K, R), blk in cs.items():
        c[I] += np.linalg.norm(blk)
    return c


def g_block_norms(sys, G, pairs, Rs):
    """||G_IJ(R)||_F for every output block."""
    from .rpa import Basis
    ao = Basis(sys.nao)
    return {(I, J, R): float(np.linalg.norm(G[R][ao.sl(I), ao.sl(J)])) for (I, J) in pairs for R in Rs}


def plain_bound(c, gp, gm, I, J, R):
    gpx, gmx = max(gp.values()), max(gm.values())
    k = (I, J, R)
    return c[I] * c[J] * (gp[k] * gmx + gm[k] * gpx + 2.0 * gpx * gmx)
