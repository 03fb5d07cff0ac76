"""Drop-in population evaluators.

Same constructor and ``evaluate_batch`` contract as the reference's
``WaveBatchEvaluator`` (``wave_batch_evaluator.py:112-191``) and
``FullNativeDecoderEvaluator`` (``full_native_decoder.py:2937-2968``):

* ``evaluate_batch(chromosomes: np.ndarray [N, 2n]) -> List[float]`` -- a
  Python list, because the BRKGA loop concatenates it with a list
  (``BRKGA_alg3.py:250``);
* infeasible chromosomes get ``1e16`` (``placement.py:9``);
* ``collision_backend`` is accepted and ignored (the native decoder ignores it
  too, ``full_native_decoder.py:2950``);
* setting ``_collect_contexts = True`` fills ``_placement_log`` with
  ``machine -> [(part_id, bin_idx, col, row, rot, shape), ...]`` as the
  reference does (``wave_batch_evaluator.py:124-136,704-706``), and
  ``_last_contexts`` with ``machine -> [BatchPlacementContext per solution]``
  whose ``open_bins`` carry each bin's grid, free runs, area, processing
  times and parts (``wave_batch_evaluator.py:74-106,251-252``), rebuilt on the
  host from the decoder's placement log (``contexts.py``).

Keys are decoded in the caller's dtype (float32 or float64), as the
placement procedure does (``placement.py:109-133``).
"""
from typing import List

import numpy as np

from .contexts import build_contexts
from .native import Decoder
from .problem import pack_problem


def _device_index(device) -> int:
    if device is None:
        return 0
    if isinstance(device, int):
        return device
    s = str(device)
    if ":" in s:
        return int(s.split(":")[1])
    return 0


class FullNativeDecoderEvaluator:
    """Population decoder on one MI355X (gfx950)."""

    def __init__(self, problem_data, nbParts, nbMachines, thresholds, instance_parts,
                 collision_backend=None, device="cuda", **decoder_options):
        self.problem_data = problem_data
        self.nbParts = int(nbParts)
        self.nbMachines = int(nbMachines)
        self.thresholds = list(thresholds)
        self.instance_parts = np.asarray(instance_parts)
        self.collision_backend = collision_backend
        self.device = device
        self.packed = pack_problem(problem_data, self.nbParts, self.nbMachines,
                                   self.thresholds, self.instance_parts)
        self._decoder = Decoder(self.packed, device=_device_index(device), **decoder_options)
        self._collect_contexts = False
        self._last_contexts = {}
        self._placement_log = {}

    def evaluate_batch(self, chromosomes) -> List[float]:
        keys = np.asarray(chromosomes)
        if keys.ndim != 2 or keys.shape[1] != 2 * self.nbParts:
            raise RuntimeError("chromosomes must have shape [N, %d], got %s"
                               % (2 * self.nbParts, keys.shape))
        if keys.dtype not in (np.float32, np.float64):
            keys = keys.astype(np.float32)
        ms, log = self._decoder.evaluate(keys, with_log=self._collect_contexts)
        if self._collect_contexts:
            self._record_log(log)
            self._last_contexts.update(build_contexts(
                self.problem_data, self.nbParts, self.nbMachines, self.thresholds,
                self.instance_parts, keys, log))
        return ms.tolist()

    def evaluate_device(self, keys, out=None, log=None, stream=None):
        """Decode device-resident keys (torch tensor [N, 2n]); returns a float64 tensor."""
        import torch
        if out is None:
            out = torch.empty(keys.shape[0], dtype=torch.float64, device=keys.device)
        return self._decoder.evaluate_device(keys, out, log=log, stream=stream)

    def _record_log(self, log):
        parts = self.problem_data.parts
        for row in log.reshape(-1, 6):
            m, pid, b, col, r, rot = (int(v) for v in row)
            if m < 0:
                continue
            shape = parts[pid].shapes[rot]
            self._placement_log.setdefault(m, []).append((pid, b, col, r, rot, shape))

    def get_stats(self):
        """Decoder counters (calls, tests, MFMA work, kernel time...)."""
        return self._decoder.stats()

    def reset_stats(self):
        self._decoder.reset_stats()

    def get_bin_stats(self):
        """Per-machine bin statistics, as the reference's native decoder reports
        them (``full_native_decoder.py:2286-2301``)."""
        return self._decoder.bin_stats()

    def reset_bin_stats(self):
        """``full_native_decoder.py:2303-2312`` (resets every counter)."""
        self._decoder.reset_stats()

    def close(self):
        self._decoder.close()


class WaveBatchEvaluator(FullNativeDecoderEvaluator):
    """Same decoder under the reference's wave-batch evaluator name."""
