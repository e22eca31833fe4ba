"""Gamma-event driver: python -m crysplight_amd.shoot_gammas <campaign.toml> [more...]
(scripts/shoot_gammas.jl): 511 keV gammas -> deposits -> optical transport on the MI355X
-> output/<tag>/events.h5."""
import sys

from .config import run_gamma_campaign


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        print("usage: python -m crysplight_amd.shoot_gammas <campaign.toml> [more...]",
              file=sys.stderr)
        return 2
    for path in argv:
        run_gamma_campaign(path)
    return 0


if __name__ == "__main__":
    sys.exit(main())
