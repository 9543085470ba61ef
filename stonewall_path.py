"""Puts the package directory jeju-stonewall-lbm_amd/ on sys.path so that
`import stonewall3d` works from the repository root (the directory name is not a valid
Python identifier)."""
import os
import sys

PKG_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "jeju-stonewall-lbm_amd")
if PKG_DIR not in sys.path:
    sys.path.insert(0, PKG_DIR)
