"""python -m rtk_amd.rtkfdk: the rtkfdk application on the HIP back end."""
import sys

from .apps import rtkfdk

if __name__ == "__main__":
    sys.exit(rtkfdk())
