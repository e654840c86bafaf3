"""``python -m ohmnivore_amd circuit.spice [--stats]``: CSV results on stdout."""
import sys

from .cli import main

sys.exit(main())
