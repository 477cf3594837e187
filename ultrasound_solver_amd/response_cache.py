"""On-disk store of exact target responses, keyed by scene and implementation.

The file is the reference's ``target_response_cache`` npz (arrays
``fingerprint``, ``points_m`` (P, 3) and ``response`` (P, M) complex;
src/algorithm/solvers/helmholtz_solver.py:146-177), so a path configured for
the reference works here.  A file whose key differs, whose arrays are
malformed or non-finite, or that cannot be read is treated as empty and is
replaced on the next store.  Points match when every coordinate agrees to
1e-12 m.
"""
import os
import tempfile
from pathlib import Path

import numpy as np

MATCH_TOLERANCE = 1.0e-12


class ResponseCache:
    def __init__(self, path, key, channels):
        self.path = Path(path) if path else None
        self.key = str(key)
        self.channels = int(channels)
        self.points = np.empty((0, 3))
        self.rows = np.empty((0, self.channels), dtype=np.complex128)
        self._load()

    def _load(self):
        if self.path is None or not self.path.is_file():
            return
        try:
            with np.load(self.path, allow_pickle=False) as data:
                if str(data["fingerprint"]) != self.key:
                    return
                points = np.array(data["points_m"], dtype=np.float64)
                rows = np.array(data["response"], dtype=np.complex128)
        except (KeyError, OSError, ValueError):
            return
        if (points.ndim == 2 and points.shape[1] == 3 and rows.shape == (len(points), self.channels)
                and np.isfinite(points).all() and np.isfinite(rows).all()):
            self.points, self.rows = points, rows

    def lookup(self, points):
        """(index of the first stored match per point, -1 where none)."""
        points = np.asarray(points, dtype=np.float64).reshape(-1, 3)
        if not len(self.points) or not len(points):
            return np.full(len(points), -1, dtype=np.int64)
        gap = np.abs(points[:, None, :] - self.points[None, :, :]).max(axis=2)
        close = gap <= MATCH_TOLERANCE
        return np.where(close.any(axis=1), close.argmax(axis=1), -1)

    def store(self, points, rows):
        """Append (points, rows) and rewrite the file atomically (a temporary
        file in the same folder, then a rename); no file without a path."""
        self.points = np.vstack((self.points, np.asarray(points, dtype=np.float64).reshape(-1, 3)))
        self.rows = np.vstack((self.rows, np.asarray(rows, dtype=np.complex128).reshape(-1, self.channels)))
        if self.path is None:
            return
        self.path.parent.mkdir(parents=True, exist_ok=True)
        fd, tmp = tempfile.mkstemp(prefix=self.path.name + ".", suffix=".npz", dir=self.path.parent)
        try:
            with os.fdopen(fd, "wb") as fh:
                np.savez_compressed(fh, fingerprint=np.asarray(self.key), points_m=self.points, response=self.rows)
            os.replace(tmp, self.path)
        except BaseException:
            Path(tmp).unlink(missing_ok=True)
            raise
