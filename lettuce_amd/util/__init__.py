"""Setup and analysis utilities (lettuce/util/utility.py, lettuce/_flow.py:
271-367) and training-data snapshots (lettuce/util/datautils.py).  They
prepare initial states, post-process fields or record them; none of them is
part of the time step."""
from .utility import *  # noqa: F401,F403
from .datautils import *  # noqa: F401,F403
