"""FP64 NumPy restatement of the gradient-surrogate MLP -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` may import this module; it checks the HIP forward, backward
and Adam kernels of ``libedhmc.so``.  The surrogate has no counterpart in the
reference (Edward differentiates the log joint, ``edward/inferences/hmc.py:199``),
so it restates published methods:

  * tanh-GELU, GELU(x) = x/2 (1 + tanh(sqrt(2/pi)(x + 0.044715 x^3)))
    (Hendrycks & Gimpel 2016), and its derivative;
  * MLP D -> H -> H -> D with PyTorch's Linear weight layout [out, in], whose
    output layer also reads the normalised input (linear skip):
    out = W3 [GELU(a2) | x] + b3 with W3 [D, H + D];
  * loss = mean over batch and coordinates of (out - (g - out_shift)/out_scale)^2
    with inputs x = (w - in_shift) * in_scale;
  * Adam (Kingma & Ba 2015) with bias correction.

``predict_bf16`` is ``predict`` with the roundings the GPU kernels apply
(bf16 inputs, weights and hidden activations on the matrix cores, fp32 bias,
scale and shift), so that a trajectory the surrogate drives can be followed
to fp32 precision instead of bf16 precision.

Pinning: the backward pass is checked against central finite differences of
the forward loss (tests/test_mlp_oracle.py), GELU' against a finite difference
of GELU, and Adam against its first two steps written out by hand;
``bf16_round`` against hand-worked ties and the float32 -> bf16 truncation
bound, ``predict_bf16`` against ``predict`` (equal when every rounded value
is already a bf16 number, within the bf16 bound otherwise).
"""

import numpy as np

K0 = np.sqrt(2.0 / np.pi)
K1 = 0.044715


def gelu(x):
  return 0.5 * x * (1.0 + np.tanh(K0 * (x + K1 * x ** 3)))


def gelu_grad(x):
  t = np.tanh(K0 * (x + K1 * x ** 3))
  return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * K0 * (1.0 + 3.0 * K1 * x * x)


def unflatten(p, D, H):
  p = np.asarray(p, np.float64)
  shapes = [('W1', (H, D)), ('b1', (H,)), ('W2', (H, H)), ('b2', (H,)), ('W3', (D, H + D)),
            ('b3', (D,))]
  out, o = {}, 0
  for k, s in shapes:
    n = int(np.prod(s))
    out[k] = p[o:o + n].reshape(s)
    o += n
  assert o == p.size
  return out


def flatten(d):
  return np.concatenate([d[k].ravel() for k in ('W1', 'b1', 'W2', 'b2', 'W3', 'b3')])


def forward(params, W, norm):
  x = (np.asarray(W, np.float64) - norm['in_shift']) * norm['in_scale']
  a1 = x @ params['W1'].T + params['b1']
  h1 = gelu(a1)
  a2 = h1 @ params['W2'].T + params['b2']
  h2 = gelu(a2)
  cat = np.concatenate([h2, x], axis=1)
  o = cat @ params['W3'].T + params['b3']
  return dict(x=x, a1=a1, h1=h1, a2=a2, h2=h2, cat=cat, o=o)


def predict(params, W, norm):
  """Surrogate gradient = out * out_scale + out_shift."""
  return forward(params, W, norm)['o'] * norm['out_scale'] + norm['out_shift']


def loss_and_grad(params, W, G, norm):
  f = forward(params, W, norm)
  t = (np.asarray(G, np.float64) - norm['out_shift']) / norm['out_scale']
  B, D = t.shape
  d = f['o'] - t
  loss = (d ** 2).sum() / (B * D)
  dO = 2.0 * d / (B * D)
  g = {}
  g['W3'] = dO.T @ f['cat']
  g['b3'] = dO.sum(0)
  H = f['h2'].shape[1]
  da2 = (dO @ params['W3'][:, :H]) * gelu_grad(f['a2'])
  g['W2'] = da2.T @ f['h1']
  g['b2'] = da2.sum(0)
  da1 = (da2 @ params['W2']) * gelu_grad(f['a1'])
  g['W1'] = da1.T @ f['x']
  g['b1'] = da1.sum(0)
  return loss, g


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
  m = b1 * m + (1 - b1) * g
  v = b2 * v + (1 - b2) * g * g
  mh = m / (1 - b1 ** t)
  vh = v / (1 - b2 ** t)
  return p - lr * mh / (np.sqrt(vh) + eps), m, v


def bf16_round(a):
  """Nearest bf16 number (ties to even) of the float32 value of a, as float64."""
  a = np.ascontiguousarray(np.asarray(a, np.float64).astype(np.float32))
  u = a.view(np.uint32).astype(np.uint64)
  u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
  return u.astype(np.uint32).view(np.float32).astype(np.float64)


def predict_bf16(params, W, norm):
  """The surrogate gradient as the GPU computes it: x = bf16(fl32((w - shift)
  * scale)) in float32 arithmetic, every layer's weights rounded to bf16,
  products accumulated exactly (FP64 here, fp32 on the matrix cores), fp32
  bias added, hidden activations rounded to bf16 after GELU, output scaled
  and shifted."""
  f32 = np.float32
  x = bf16_round((np.asarray(W, np.float64).astype(f32) - np.asarray(norm['in_shift'], f32))
                 * np.asarray(norm['in_scale'], f32))
  h1 = bf16_round(gelu(x @ bf16_round(params['W1']).T + params['b1']))
  h2 = bf16_round(gelu(h1 @ bf16_round(params['W2']).T + params['b2']))
  o = np.concatenate([h2, x], axis=1) @ bf16_round(params['W3']).T + params['b3']
  o = o.astype(f32)
  return (o * np.asarray(norm['out_scale'], f32) + np.asarray(norm['out_shift'], f32)
          ).astype(np.float64)
