"""TOML front end: ``run(input_file, pes=...)`` (src/pyticc/input/driver.py of the reference)."""

from pyticc_amd.input.driver import run

__all__ = ["run"]
