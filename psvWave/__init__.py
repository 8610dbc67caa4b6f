"""The reference's import name.

Code written against psvWave (``import psvWave``; psvWave/__init__.py:6,
34-125) runs unchanged: ``psvWave.fdModel``, ``get_dictionary``,
``write_dictionary`` and ``from_dictionary`` are those of ``psvwave_amd``,
which runs the fdModel hot path on the MI355X.
"""
from psvwave_amd import (Batch, Config, __version__, fdModel, from_dictionary,
                         get_dictionary, write_dictionary)

__all__ = ["fdModel", "get_dictionary", "write_dictionary", "from_dictionary",
           "Config", "Batch"]
