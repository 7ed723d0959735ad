"""``python -m darcy``: the command line of darcy.cli."""

from darcy.cli import main

raise SystemExit(main())
