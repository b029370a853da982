"""Global configuration (frostspin/config.py:1-37).

`config["SU2_CG_file"]` names a Clebsch-Gordan table in the reference's JSON or
npz format, from the command-line option `--frostspin-SU2-CG-file` like the
reference. The coefficients are computed exactly on demand here, so the option
is not required: without it no table is read, and a table that is given takes
precedence over the computed values for the dimensions it holds (a missing file
is an error, as in the reference). `config["debug_checks"]` switches the norm
and structure checks that run inside assert statements (on unless python -O).
`config["ctmrg_graphs"]` lets SequentialCTMRG record settled iterations as HIP
graphs and replay them (frostspin_amd.ctmrg.graphs; GPU only, on by default).
"""

import argparse

__version__ = "1.0.0"

ASSERT_TOL = 4e-13

_parser = argparse.ArgumentParser(add_help=False)
_parser.add_argument(
    "--frostspin-SU2-CG-file",
    help="SU(2) Clebsch-Gordan table (json or npz); coefficients are computed when absent",
    type=str,
    default=None,
)
_args, _ = _parser.parse_known_args()

config = {"SU2_CG_file": _args.frostspin_SU2_CG_file, "debug_checks": __debug__, "ctmrg_graphs": True}

if not __debug__:
    print("\nInfo: assert statements are disabled")
