#!/usr/bin/env python3
"""Launch-interface generator command line with the reference's argv (see generator/cli.py)."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.realpath(__file__))))
from threadcoarsenedstencilgenerator_amd.generator.cli import generate_launch  # noqa: E402

if __name__ == "__main__":
    sys.exit(generate_launch())
