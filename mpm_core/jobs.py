"""What the bake runner and the surfacer share around their frame loops.

Both are jobs on one folder of a cache (the reference's runners, src/mpm_core/bake.py:41-53, 116-145, 211-244
and src/mpm_surface/run.py:40-75, src/mpm_surface/__main__.py:1-42, behave alike here):

* one run per folder, held by an operating-system lock file; a second run gets RuntimeError;
* a ``cancel`` file in the folder stops the run after the current frame; one left over from an earlier run is
  removed when a run starts;
* ``progress.json`` (status, frame, frame_end, particles, ms_per_frame, message, warnings), reset to
  "starting" as soon as the run holds the lock and set to "error" with the exception when one escapes;
* a command line ``python -m <module> CACHE_DIR [--device D] [--watch-stdin] ...`` with exit status 0 when
  done or cancelled and 1 on error.
"""

from __future__ import annotations

import argparse
import sys
from collections.abc import Callable
from contextlib import ExitStack, contextmanager
from pathlib import Path

from mpm_core.cache import exclusive_lock, stdin_closed, write_json


class ProgressFile:
    """A progress.json, rewritten whole on every update. Without a path the document is only kept in memory
    (a bake server request without a cache folder)."""

    def __init__(self, path: Path | None = None):
        self.path = None if path is None else Path(path)
        self.doc = dict(status="starting", frame=None, frame_end=None, particles=0, ms_per_frame=None, message="",
                        warnings=[])

    def warn(self, message: str) -> None:
        if message not in self.doc["warnings"]:
            self.doc["warnings"].append(message)

    def update(self, **fields) -> None:
        self.doc.update(fields)
        if self.path is not None:
            write_json(self.path, self.doc)


class FolderJob:
    """One run on ``folder``: ``with FolderJob(...) as job:`` holds the lock, clears a stale cancel file and
    publishes the "starting" progress; ``job.recording_errors()`` wraps the run's body so that an escaping
    exception is recorded in progress.json before it propagates."""

    def __init__(self, folder: Path, lock_name: str, busy_message: str,
                 should_cancel: Callable[[], bool] | None = None):
        self.folder = Path(folder)
        self.lock_path = self.folder / lock_name
        self.busy_message = busy_message
        self.should_cancel = should_cancel
        self.cancel_file = self.folder / "cancel"
        self.progress = ProgressFile(self.folder / "progress.json")
        self._held = ExitStack()

    def __enter__(self) -> FolderJob:
        try:
            self._held.enter_context(exclusive_lock(self.lock_path))
        except RuntimeError as exc:
            raise RuntimeError(self.busy_message) from exc
        self.cancel_file.unlink(missing_ok=True)
        self.progress.update()
        return self

    def __exit__(self, *exc) -> None:
        self._held.close()

    def cancel_requested(self) -> bool:
        """A cancel file (consumed here) or, failing that, the caller's callback."""
        if self.cancel_file.exists():
            self.cancel_file.unlink()
            return True
        return bool(self.should_cancel and self.should_cancel())

    @contextmanager
    def recording_errors(self):
        """Around the run's body: an exception that escapes is written to progress.json, then re-raised."""
        try:
            yield self
        except Exception as exc:
            self.progress.update(status="error", message=f"{type(exc).__name__}: {exc}")
            raise


def command_line(prog: str, doc: str, argv: list[str] | None, add_arguments: Callable[[argparse.ArgumentParser], None],
                 run: Callable[[argparse.Namespace, Callable[[], bool] | None], object],
                 summary: Callable[[object], str]) -> int:
    """Parse ``CACHE_DIR [--device D] [--watch-stdin]`` plus the job's own options, run the job and print its
    summary line. ``run`` gets the parsed arguments and a callable that turns true once standard input has
    closed (with --watch-stdin: the launching process has gone, which cancels the job), else None. Exit
    status 1 on an exception (printed to stderr) or an "error" result, else 0."""
    ap = argparse.ArgumentParser(prog=prog, description=doc.split("\n\n")[0])
    ap.add_argument("cache_dir", type=Path)
    ap.add_argument("--device", default=None)
    add_arguments(ap)
    ap.add_argument("--watch-stdin", action="store_true",
                    help="stop (as if cancelled) when stdin closes, i.e. when the launching process exits")
    args = ap.parse_args(argv)
    parent_gone = stdin_closed().is_set if args.watch_stdin else None
    try:
        result = run(args, parent_gone)
    except Exception as exc:
        print(f"error: {type(exc).__name__}: {exc}", file=sys.stderr)
        return 1
    print(summary(result))
    return 1 if getattr(result, "status", None) == "error" else 0
