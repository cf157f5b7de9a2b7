"""`python -m amls_amd.solve`: the amls_solve driver (src/apps/amls_solve.cpp)."""
import sys

from .apps import solve_main

if __name__ == "__main__":
    sys.exit(solve_main())
