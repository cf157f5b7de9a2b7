"""`python -m amls_amd.verify`: the amls_verify driver (src/apps/amls_verify.cpp)."""
import sys

from .apps import verify_main

if __name__ == "__main__":
    sys.exit(verify_main())
