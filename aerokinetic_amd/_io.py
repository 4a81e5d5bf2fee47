"""Atomic local file output (temporary file + fsync + rename)."""
from contextlib import contextmanager
import hashlib
import json
import os
from pathlib import Path
import tempfile


@contextmanager
def atomic_binary(path):
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    fd, tmp = tempfile.mkstemp(prefix="." + path.name + ".", dir=path.parent)
    try:
        with os.fdopen(fd, "wb") as handle:
            yield handle
            handle.flush()
            os.fsync(handle.fileno())
        os.replace(tmp, path)
    finally:
        if os.path.exists(tmp):
            os.unlink(tmp)


def write_json(path, data):
    text = json.dumps(data, ensure_ascii=False, indent=2, allow_nan=False) + "\n"
    with atomic_binary(path) as handle:
        handle.write(text.encode("utf-8"))


def file_hash(path, limit=None):
    """sha256 of a file, or of its first `limit` bytes (which must exist)."""
    digest = hashlib.sha256()
    remaining = limit
    with Path(path).open("rb") as handle:
        while remaining is None or remaining > 0:
            size = 1 << 20 if remaining is None else min(remaining, 1 << 20)
            chunk = handle.read(size)
            if not chunk:
                break
            digest.update(chunk)
            if remaining is not None:
                remaining -= len(chunk)
    if remaining:
        raise ValueError("History is shorter than the saved checkpoint")
    return digest.hexdigest()
