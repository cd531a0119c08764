#!/usr/bin/env python3
"""Kernel generator command line with the reference's argv (see generator/cli.py)."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.realpath(__file__))))
from threadcoarsenedstencilgenerator_amd.generator.cli import generate  # noqa: E402

if __name__ == "__main__":
    sys.exit(generate(2))
