"""Surface a baked particle cache.

    python -m mpm_surface CACHE_DIR [--frames A-B] [--device cuda:0] [--watch-stdin]

Settings come from CACHE_DIR/surface_settings.json (the defaults where it is missing). Exit status 0 when
done or cancelled, 1 on error (the reference's command line, src/mpm_surface/__main__.py:1-42).
"""

import sys

from mpm_core.jobs import command_line
from mpm_surface.run import surface


def frame_range(text: str) -> list[int]:
    """``"A-B"``: frames A to B inclusive; ``"A"``: frame A alone."""
    bounds = [int(part) for part in text.split("-", 1)]
    return list(range(bounds[0], bounds[-1] + 1))


def main(argv: list[str] | None = None) -> int:
    return command_line(
        "python -m mpm_surface", __doc__, argv,
        lambda ap: ap.add_argument("--frames", type=frame_range, default=None, help="e.g. 1-48"),
        lambda args, parent_gone: surface(args.cache_dir, frames=args.frames, device=args.device,
                                          should_cancel=parent_gone),
        lambda r: f"{r.status}: {r.frames_written} frames meshed, {r.frames_skipped} unchanged")


if __name__ == "__main__":
    sys.exit(main())
