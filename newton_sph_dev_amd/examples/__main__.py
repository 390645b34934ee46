"""Entry point: python -m newton_sph_dev_amd.examples <example_name> [options]."""

from . import main

if __name__ == "__main__":
    main()
