"""python -m schnellfmm_amd: the lfmm2 command line (src/main.rs:12-289)."""
import sys

from .cli import main

sys.exit(main())
