"""Point-source driver: python -m crysplight_amd.run_point_source <config.toml> [more...]
(scripts/run_point_source.jl)."""
import sys

from .config import run_from_config


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        print("usage: python -m crysplight_amd.run_point_source <config.toml> [more configs...]",
              file=sys.stderr)
        return 2
    for path in argv:
        run_from_config(path)
    return 0


if __name__ == "__main__":
    sys.exit(main())
