"""The tests' handle on the CPU oracle (the package under oracle/)."""
from oracle import *  # noqa: F401,F403
from oracle import Oracle, load_oracle  # noqa: F401
