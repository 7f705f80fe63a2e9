"""Operators, flows and reporters (lettuce.ext counterparts)."""
from ._operators import *  # noqa: F401,F403
from ._force import *  # noqa: F401,F403
from ._learned import *  # noqa: F401,F403
from ._training import *  # noqa: F401,F403
from ._flows import *  # noqa: F401,F403
from ._reporter import *  # noqa: F401,F403
from ._failure import *  # noqa: F401,F403
from ._slab import *  # noqa: F401,F403
