"""Host-side assembly of multi-beam work: beamlet/PTV flags of a 4pi plan on the device, and every
beamlet of a set of beams concatenated once so that requests of any beams are cheap to build.

The reference builds its beams inside IMRT::BeamBundleInit / beamletFlagInit / beamletsInit
(IMRTOpt/src/IMRTBeamBundle.cu:15-127, 130-316, 319-569) and hands them to DoseMatConstruction one
beam at a time; these helpers feed the same beams to the batched and serving entry points.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import host
from ._lib import Beam


def device_flags(ctx, cfg, torch):
    """beamletFlagInit of every bundle of a 4pi config on the device (one call per PTV segment, since
    each bundle is flagged against its own segment's mask)."""
    segs, seg_of = cfg.notes["segments"], cfg.notes["segment_of_beam"]
    flags = [None] * len(cfg.bundles)
    for s_idx, mask in enumerate(segs):
        ks = [k for k in range(len(cfg.bundles)) if seg_of[k] == s_idx]
        if not ks:
            continue
        mask_d = torch.from_numpy(np.ascontiguousarray(mask, dtype=np.uint8)).cuda()
        for k, f in zip(ks, ctx.beamlet_flags(mask_d, [cfg.bundles[k] for k in ks])):
            flags[k] = f
    return flags


class BeamStore:
    """All beamlets of a set of beams (beamletsInit of each bundle with its flags), concatenated."""

    def __init__(self, cfg, flags):
        per = [host.beamlets_init(b, f, cfg.volume) for b, f in zip(cfg.bundles, flags)]
        self.counts = np.array([len(b) for b, _ in per], dtype=np.int64)
        self.first = np.concatenate([[0], np.cumsum(self.counts)]).astype(np.int64)
        n = int(self.first[-1])
        self.n = n
        self.beams = (Beam * max(n, 1))()
        for k, (b, _) in enumerate(per):
            if len(b):
                C.memmove(C.byref(self.beams, int(self.first[k]) * C.sizeof(Beam)), b, C.sizeof(Beam) * len(b))
        self.flu = np.concatenate([f for _, f in per], 0) if n else np.zeros((0, 256), np.float32)

    def beamlet(self, g: int) -> Beam:
        """Beamlet g of the concatenation (a copy)."""
        b = Beam()
        C.memmove(C.byref(b), C.byref(self.beams, g * C.sizeof(Beam)), C.sizeof(Beam))
        return b

    def beam_of(self, g: int):
        """(beam k, row within beam k) of global beamlet g."""
        k = int(np.searchsorted(self.first, g, side="right") - 1)
        return k, g - int(self.first[k])

    def request(self, ks):
        """Beamlets (ctypes array), fluence (n x 256) and per-beam counts of beams ks, in that order."""
        ks = list(ks)
        cnt = [int(self.counts[k]) for k in ks]
        arr = (Beam * max(sum(cnt), 1))()
        o = 0
        for k, c in zip(ks, cnt):
            C.memmove(C.byref(arr, o * C.sizeof(Beam)), C.byref(self.beams, int(self.first[k]) * C.sizeof(Beam)),
                      C.sizeof(Beam) * c)
            o += c
        flu = np.concatenate([self.flu[self.first[k]:self.first[k] + self.counts[k]] for k in ks], 0) \
            if ks else np.zeros((0, 256), np.float32)
        return arr, flu, cnt
