"""Crash-safe local output of a run: whole files replaced in one rename, and append-only CSV
logs whose durable prefix a checkpoint can pin.

* `PublishedFile(path)` / `publish(path, data)`: the bytes go to a temporary sibling that is
  flushed, fsync'ed and renamed over `path`, then the directory entry is synced, so a reader
  or a crash sees the old file or the new one, never a torn mix.
* `write_json`: the run files' JSON text (two-space indent, UTF-8 kept, NaN refused, final
  newline), as the reference writes metadata.json / summary.json / project.json.
* `prefix_digest(path, nbytes)`: SHA-256 of the first `nbytes` bytes (of all of them for None).
* `CsvLog`: history.csv / residuals.csv.  `durable()` syncs it and returns its size and
  digest for a checkpoint; a resumed run reopens it at that prefix, refusing a file whose
  prefix changed, and moves anything written after it to recovery/<token>/.
"""
from __future__ import annotations

import csv
import hashlib
import json
import os
from pathlib import Path
import shutil
import tempfile
import uuid


def _sync_directory(directory):
    try:
        fd = os.open(directory, os.O_RDONLY)
    except OSError:
        return
    try:
        os.fsync(fd)
    except OSError:
        pass
    finally:
        os.close(fd)


class PublishedFile:
    """Binary handle whose bytes replace `path` when the block exits without an exception
    (on an exception the partial file is removed and `path` is left as it was)."""

    def __init__(self, path):
        self.path = Path(path)
        self._temp = None
        self._handle = None

    def __enter__(self):
        self.path.parent.mkdir(parents=True, exist_ok=True)
        fd, self._temp = tempfile.mkstemp(dir=self.path.parent, prefix=f".{self.path.name}.", suffix=".part")
        self._handle = os.fdopen(fd, "wb")
        return self._handle

    def __exit__(self, kind, value, traceback):
        try:
            if kind is None:
                self._handle.flush()
                os.fsync(self._handle.fileno())
            self._handle.close()
            if kind is None:
                os.replace(self._temp, self.path)
                _sync_directory(self.path.parent)
        finally:
            if os.path.exists(self._temp):
                os.unlink(self._temp)
        return False


def publish(path, data: bytes):
    with PublishedFile(path) as handle:
        handle.write(data)


def write_json(path, data):
    publish(path, (json.dumps(data, ensure_ascii=False, indent=2, allow_nan=False) + "\n").encode("utf-8"))


def prefix_digest(path, nbytes=None):
    """SHA-256 hex digest of the first `nbytes` bytes of `path` (the whole file for None)."""
    digest = hashlib.sha256()
    buffer = bytearray(1 << 20)
    view = memoryview(buffer)
    left = nbytes
    with open(path, "rb", buffering=0) as f:
        while left is None or left > 0:
            got = f.readinto(view if left is None else view[:min(left, len(buffer))])
            if not got:
                break
            digest.update(view[:got])
            if left is not None:
                left -= got
    if left:
        raise ValueError(f"{Path(path).name} holds fewer bytes than the checkpoint recorded")
    return digest.hexdigest()


class Recovery:
    """recovery/<token>/ of one resumed run: created on first use, one token per resume."""

    def __init__(self, output):
        self.output = Path(output)
        self._dir = None

    def directory(self):
        if self._dir is None:
            self._dir = self.output / "recovery" / uuid.uuid4().hex
            self._dir.mkdir(parents=True)
        return self._dir


class CsvLog:
    """Append-only CSV log of a run.  `resume` = (bytes, sha256) of the prefix a checkpoint
    recorded; rows after it are kept as recovery/<token>/<stem>_before_resume.csv."""

    def __init__(self, path, columns, *, resume=None, recovery: Recovery | None = None):
        self.path = Path(path)
        self.columns = list(columns)
        if resume is None:
            self._file = self.path.open("w", newline="", encoding="utf-8")
            self._rows = csv.DictWriter(self._file, self.columns)
            self._rows.writeheader()
            self._file.flush()
            return
        size, digest = resume
        if prefix_digest(self.path, size) != digest:
            raise ValueError(f"{self.path.name} before the checkpoint has changed; refusing to mix results")
        if self.path.stat().st_size > size:
            keep = (recovery or Recovery(self.path.parent)).directory() / f"{self.path.stem}_before_resume.csv"
            shutil.copy2(self.path, keep)
            os.truncate(self.path, size)
        self._file = self.path.open("a", newline="", encoding="utf-8")
        self._rows = csv.DictWriter(self._file, self.columns)

    def read(self):
        """The rows written so far (as strings)."""
        self._file.flush()
        with self.path.open(newline="", encoding="utf-8") as f:
            return list(csv.DictReader(f))

    def append(self, row):
        self._rows.writerow({k: row.get(k, "") for k in self.columns})

    def flush(self):
        self._file.flush()

    def durable(self):
        """Sync the log; returns (bytes, sha256) of its content, for a checkpoint."""
        self._file.flush()
        os.fsync(self._file.fileno())
        size = self.path.stat().st_size
        return size, prefix_digest(self.path, size)

    def close(self):
        if not self._file.closed:
            self._file.close()
