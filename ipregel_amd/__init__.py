"""MI355X (gfx950) engine for iPregel's superstep hot path.

Native parts (built in-tree by ipregel_amd/build.py):
  lib/libipregel_hip.so       thin HIP C-ABI (include/ipregel_hip.h)
  lib/libipregel_graph.so     graph generators and the .config/.idx/.adj triple
  lib/libipr_<app>_32.so      an application TU (include/iPregel.h) + C driver
  bin/<app>_32                the same application as a command-line program
Python modules: graph (generators / file triple), engine (ctypes harness).
"""
