"""Setup and analysis utilities (lettuce/util/utility.py, lettuce/_flow.py:
271-367).  They prepare initial states or post-process fields; none of them
is part of the time step."""
from .utility import *  # noqa: F401,F403
