"""python -m rtk_amd.rtkconjugategradient: the rtkconjugategradient application on the HIP back end."""
import sys

from .apps import rtkconjugategradient

if __name__ == "__main__":
    sys.exit(rtkconjugategradient())
