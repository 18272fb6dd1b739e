"""Learned gradient surrogate for HMC (no counterpart in the reference).

An MLP D -> H -> H -> D (tanh-GELU) approximates grad log p(w | data).  It is
trained on the GPU by libedhmc.so's own forward/backward/Adam kernels on
(w, grad) pairs that the plain sampler records during warm-up, then frozen and
used as the leapfrog force; the exact log joint (bf16 X, hi/lo split
positions, FP64 reduction) still decides every Metropolis test, so the chain
targets the exact posterior.  A frozen surrogate keeps the kernel a fixed,
reversible, volume-preserving proposal; adapting it during sampling would
break the Markov property.
"""

import ctypes
import math

import numpy as np
import torch

from edward_amd import _lib as L


class Surrogate(object):
  """Gradient surrogate attached to a bf16 ``GLMEngine``.

  Args:
    engine: the GLMEngine whose posterior is sampled (precision 'bf16').
    hidden: hidden width H.
    max_batch: largest batch for training / number of chains.
    seed: seed of the weight initialisation (Glorot-uniform).
  """

  def __init__(self, engine, hidden=1024, max_batch=4096, seed=0):
    self.engine = engine
    self.D = engine.D
    self.H = int(hidden)
    self.max_batch = int(max_batch)
    # the initial weights are drawn on the host before the library call,
    # whose allocation waits for the device: work queued before runs meanwhile
    flat = self._initial_params(seed)
    engine.check(engine.lib.edhmc_sur_create(engine.h, self.D, self.H, self.max_batch))
    self.n_params = int(engine.lib.edhmc_sur_num_params(engine.h))
    self.set_params(flat)
    self.step_count = 0

  # -- parameters ----------------------------------------------------------
  def init_params(self, seed=0):
    self.set_params(self._initial_params(seed))
    self.step_count = 0

  def _initial_params(self, seed):
    rng = np.random.default_rng(seed)
    D, H = self.D, self.H

    def glorot(o, i):
      a = math.sqrt(6.0 / (i + o))
      return rng.uniform(-a, a, size=(o, i))
    # output layer [hidden | skip]: small hidden part, zero skip (fit() starts
    # the skip at the least-squares linear map)
    W3 = np.concatenate([0.1 * glorot(D, H), np.zeros((D, D))], axis=1)
    return np.concatenate([glorot(H, D).ravel(), np.zeros(H), glorot(H, H).ravel(), np.zeros(H),
                           W3.ravel(), np.zeros(D)]).astype(np.float32)

  def _offsets(self):
    D, H = self.D, self.H
    o = {}
    o['W1'] = 0
    o['b1'] = o['W1'] + H * D
    o['W2'] = o['b1'] + H
    o['b2'] = o['W2'] + H * H
    o['W3'] = o['b2'] + H
    o['b3'] = o['W3'] + D * (H + D)
    return o

  def init_skip_least_squares(self, W, G, rows=None, max_rows=32768, ridge=1e-6):
    """Start the output layer's linear skip at the least-squares map from the
    normalised inputs to the normalised targets, over an evenly spaced sample
    of ``rows`` (default all) of the pairs (W, G): the Gram products in FP64
    on the device (``edhmc_sur_skip_gram``), the D x D solve on the host."""
    eng = self.engine
    P = int(W.shape[0])
    rows = np.arange(P) if rows is None else np.asarray(rows, np.int64)
    sel = rows[np.linspace(0, len(rows) - 1, min(len(rows), max_rows)).astype(np.int64)]
    idx = torch.as_tensor(sel, dtype=torch.int64).to(eng.device)
    xtx = torch.empty(self.D, self.D, dtype=torch.float64, device=eng.device)
    xtt = torch.empty_like(xtx)
    eng.check(eng.lib.edhmc_sur_skip_gram(eng.h, L.ptr(W), L.ptr(G), P, L.ptr(idx), len(sel),
                                          L.ptr(xtx), L.ptr(xtt), eng.stream))
    xtx = xtx.cpu().numpy()
    xtt = xtt.cpu().numpy()
    A = np.linalg.solve(xtx + ridge * len(sel) * np.eye(self.D), xtt)   # t ~ x A
    # written into the device copy of the parameters (only the D x D block
    # crosses to the device)
    flat = self.get_params()
    o = self._offsets()
    W3 = flat[o['W3']:o['b3']].view(self.D, self.H + self.D)
    W3[:, self.H:] = torch.from_numpy(np.ascontiguousarray(A.T, dtype=np.float32)).to(eng.device)
    flat[o['b3']:o['b3'] + self.D] = 0.0
    self.set_params(flat)
    self.step_count = 0

  def set_params(self, flat):
    if isinstance(flat, torch.Tensor):
      p = flat.to(self.engine.device, torch.float32).contiguous()
    else:
      p = torch.as_tensor(np.asarray(flat, np.float32)).to(self.engine.device)
    if p.numel() != self.n_params:
      raise ValueError('expected %d parameters, got %d' % (self.n_params, p.numel()))
    self.engine.check(self.engine.lib.edhmc_sur_set_params(self.engine.h, L.ptr(p),
                                                           self.engine.stream))
    torch.cuda.synchronize(self.engine.device)

  def get_params(self):
    p = torch.empty(self.n_params, device=self.engine.device)
    self.engine.check(self.engine.lib.edhmc_sur_get_params(self.engine.h, L.ptr(p),
                                                           self.engine.stream))
    return p

  def set_normalisation(self, in_shift, in_scale, out_shift, out_scale):
    dev = self.engine.device
    vs = [torch.as_tensor(v, dtype=torch.float32).to(dev).contiguous().reshape(-1)
          for v in (in_shift, in_scale, out_shift, out_scale)]
    self._norm = vs
    self.engine.check(self.engine.lib.edhmc_sur_set_norm(
        self.engine.h, *[L.ptr(v) for v in vs], self.engine.stream))

  def normalise_from(self, W, G):
    """Per-coordinate standardisation of inputs and targets from the data
    (column statistics on the device, the D-vector arithmetic on the host)."""
    wm, ws = [v.cpu().numpy().astype(np.float64) for v in self.engine.col_stats(W)]
    gm, gs = [v.cpu().numpy().astype(np.float64) for v in self.engine.col_stats(G)]
    ws = np.where(ws > 0, ws, 1.0)
    gs = np.where(gs > 0, gs, 1.0)
    self.set_normalisation(wm, 1.0 / ws, gm, gs)

  # -- compute -------------------------------------------------------------
  def forward(self, W):
    W = W.to(self.engine.device, torch.float32).contiguous()
    out = torch.empty_like(W)
    for s in range(0, int(W.shape[0]), self.max_batch):
      Wc = W[s:s + self.max_batch]
      oc = out[s:s + self.max_batch]
      self.engine.check(self.engine.lib.edhmc_sur_forward(
          self.engine.h, L.ptr(Wc), int(Wc.shape[0]), L.ptr(oc), self.engine.stream))
    return out

  def loss_grad(self, W, G):
    W = W.to(self.engine.device, torch.float32).contiguous()
    G = G.to(self.engine.device, torch.float32).contiguous()
    loss = torch.empty(1, dtype=torch.float64, device=self.engine.device)
    grad = torch.empty(self.n_params, device=self.engine.device)
    self.engine.check(self.engine.lib.edhmc_sur_loss_grad(
        self.engine.h, L.ptr(W), L.ptr(G), int(W.shape[0]), L.ptr(loss), L.ptr(grad),
        self.engine.stream))
    return loss, grad

  def train_step(self, W, G, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    self.step_count += 1
    loss = torch.empty(1, dtype=torch.float64, device=self.engine.device)
    self.engine.check(self.engine.lib.edhmc_sur_train_step(
        self.engine.h, L.ptr(W), L.ptr(G), int(W.shape[0]), float(lr), float(beta1),
        float(beta2), float(eps), self.step_count, L.ptr(loss), self.engine.stream))
    return loss

  def train_epoch(self, W, G, perm, batch, lrs, beta1=0.9, beta2=0.999, eps=1e-8):
    """len(lrs) Adam steps, step b on the pairs (W, G) at the indices
    perm[b*batch:(b+1)*batch] with learning rate lrs[b], in one library call
    (no host round trip per step).  Returns the per-step losses (device
    fp64)."""
    dev = self.engine.device
    n_batches = len(lrs)
    lr_host = (ctypes.c_float * max(n_batches, 1))(*[float(x) for x in lrs])
    losses = torch.empty(n_batches, dtype=torch.float64, device=dev)
    perm = perm.to(dev, torch.int64).contiguous()
    self.engine.check(self.engine.lib.edhmc_sur_train_epoch(
        self.engine.h, L.ptr(W), L.ptr(G), int(W.shape[0]), L.ptr(perm), int(batch),
        n_batches, lr_host, float(beta1), float(beta2), float(eps), self.step_count,
        L.ptr(losses), self.engine.stream))
    self.step_count += int(n_batches)
    return losses

  def _launch_validation(self, W, G):
    """Launch the loss of (W, G) in batches of max_batch (forward only, no
    parameter update); returns the batches' losses (device fp64, not yet
    read) and their sizes."""
    eng = self.engine
    W = W.to(eng.device, torch.float32).contiguous()
    G = G.to(eng.device, torch.float32).contiguous()
    starts = list(range(0, int(W.shape[0]), self.max_batch))
    losses = torch.empty(max(len(starts), 1), dtype=torch.float64, device=eng.device)
    sizes = []
    for k, s in enumerate(starts):
      n = min(self.max_batch, int(W.shape[0]) - s)
      eng.check(eng.lib.edhmc_sur_loss_grad(
          eng.h, L.ptr(W[s:s + n]), L.ptr(G[s:s + n]), n, L.ptr(losses[k:k + 1]), None,
          eng.stream))
      sizes.append(n)
    return losses[:len(sizes)], sizes

  @staticmethod
  def _mean_loss(batch_losses, sizes):
    if not sizes:
      return 0.0
    return float(np.dot(np.asarray(batch_losses, np.float64), np.asarray(sizes, np.float64))
                 / sum(sizes))

  def validation_loss(self, W, G):
    """Mean loss over (W, G) in batches of max_batch (no parameter update;
    the batches' losses land in one device array, read once)."""
    losses, sizes = self._launch_validation(W, G)
    return self._mean_loss(losses.cpu().numpy(), sizes)

  def fit(self, W, G, epochs=20, batch=4096, lr=1e-3, lr_final=1e-4, seed=0, normalise=True,
          log=None, val_frac=0.0):
    """Adam on the pairs (W [P, D], G [P, D], device tensors), reshuffled
    every epoch, learning rate decayed geometrically from lr to lr_final.
    With val_frac > 0 that fraction of the pairs is held out and the weights
    of the epoch with the lowest held-out loss are kept (early stopping).
    Shuffles are host permutations of row indices; each epoch's batches
    are read through them by ``edhmc_sur_train_epoch``, the held-out rows
    gathered once (``edhmc_gather_rows``).
    Returns the mean training loss of each epoch."""
    eng = self.engine
    dev = eng.device
    W = W.to(dev, torch.float32).contiguous()
    G = G.to(dev, torch.float32).contiguous()
    P_all = int(W.shape[0])
    rng = np.random.default_rng(int(seed))
    order = rng.permutation(P_all)
    nv = max(1, int(val_frac * P_all)) if val_frac > 0 else 0
    train_rows = order[nv:]
    Wv = Gv = None
    if nv:
      vi = torch.as_tensor(order[:nv], dtype=torch.int64).to(dev)
      Wv = eng.gather_rows(W, vi)
      Gv = eng.gather_rows(G, vi)
    P = len(train_rows)
    batch = min(int(batch), self.max_batch, P)
    if normalise:
      self.normalise_from(W, G)
      self.init_skip_least_squares(W, G, rows=train_rows)
    best = None
    if Wv is not None:
      best = (self.validation_loss(Wv, Gv), self.get_params().clone())
    n_batches = P // batch
    total = max(1, epochs * n_batches)
    decay = (lr_final / lr) ** (1.0 / total) if lr_final else 1.0
    losses = []
    step = 0
    perm = torch.as_tensor(train_rows[rng.permutation(P)], dtype=torch.int64).to(dev)
    for ep in range(epochs):
      lrs = [lr * decay ** (step + b) for b in range(n_batches)]
      ep_losses = self.train_epoch(W, G, perm, batch, lrs)
      # the held-out loss is queued behind the epoch, and the next epoch's
      # shuffle drawn and sent (from pinned memory, without waiting) while
      # they run; the host then reads both losses at once
      if best is not None:
        vl_dev, vsizes = self._launch_validation(Wv, Gv)
        read = torch.cat([ep_losses, vl_dev])
      else:
        read = ep_losses
      if ep + 1 < epochs:
        perm = torch.as_tensor(train_rows[rng.permutation(P)], dtype=torch.int64).pin_memory()
        perm = perm.to(dev, non_blocking=True)
      vals = read.cpu().numpy()
      ep_loss = float(vals[:n_batches].mean())
      step += n_batches
      losses.append(ep_loss)
      msg = 'surrogate epoch %d loss %.5f' % (ep, ep_loss)
      if best is not None:
        vl = self._mean_loss(vals[n_batches:], vsizes)
        msg += ' held-out %.5f' % vl
        if vl < best[0]:
          best = (vl, self.get_params().clone())
      if log:
        log(msg)
    if best is not None:
      steps = self.step_count
      self.set_params(best[1])
      self.step_count = steps
      self.best_validation_loss = best[0]
    return losses

  # -- sampling ------------------------------------------------------------
  def attach(self, engine):
    if engine is not self.engine:
      raise ValueError('the surrogate belongs to another engine')

  def transition(self, step_size, n_steps, momenta=None, uniforms=None, seed=0, counter=0,
                 sample_out=None, n_accept=None, accepted_out=None):
    eng = self.engine
    r0 = None if momenta is None else torch.as_tensor(momenta, dtype=torch.float32).to(
        eng.device).contiguous()
    u = None if uniforms is None else torch.as_tensor(uniforms, dtype=torch.float64).to(
        eng.device).contiguous()
    eng.check(eng.lib.edhmc_transition_surrogate(
        eng.h, float(step_size), int(n_steps), L.ptr(r0), L.ptr(u), int(seed), int(counter),
        L.ptr(accepted_out), L.ptr(sample_out), L.ptr(n_accept), eng.stream))
