"""Progress bar with the output format of ``edward/util/progbar.py:11-115``:
``t/T [ pct%] bar ETA: s | Acceptance Rate: x.xxx``."""

import sys
import time


class Progbar(object):

  def __init__(self, target, width=30, interval=0.01, verbose=1, stream=None):
    self.target = target
    self.width = width
    self.interval = interval
    self.verbose = verbose
    self.stream = stream or sys.stdout
    self.stored_values = {}
    self.start = time.time()
    self.last_update = 0.0
    self.total_width = 0

  def update(self, current, values=None, force=False):
    for k, v in (values or {}).items():
      self.stored_values[k] = v
    now = time.time()
    if not force and now - self.last_update < self.interval and current < self.target:
      return
    self.last_update = now
    if self.verbose == 0:
      return
    out = self.stream
    prev = self.total_width
    out.write('\b' * prev)
    out.write('\r')
    nd = len(str(self.target))
    bar = '%*d/%*d' % (nd, current, nd, self.target)
    pct = int(current / float(self.target) * 100) if self.target else 100
    bar += ' [{0}%] '.format(str(pct).rjust(3))
    fill = int(self.width * float(current) / self.target) if self.target else self.width
    bar += '#' * fill + ' ' * (self.width - fill)
    per = (now - self.start) / current if current else 0.0
    if current < self.target:
      info = ' ETA: %ds' % (per * (self.target - current))
    else:
      info = ' Elapsed: %ds' % (now - self.start)
    for k, v in self.stored_values.items():
      info += ' | {0:s}: {1:0.3f}'.format(k, v)
    self.total_width = len(bar) + len(info)
    if prev > self.total_width:
      info += ' ' * (prev - self.total_width)
    out.write(bar + info)
    out.flush()
    if current >= self.target:
      out.write('\n')
