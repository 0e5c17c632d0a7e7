"""Copy two time-history files of the reference's QA runs into tests/golden (data fixtures).

LOI14T01: /TH/BRIC (DEF) on 3 solids and /TH/NODE (DEF) on 3 nodes, 12 global variables
(qa-tests/miniqa/LOIS/LOI14/reference, deck LOI14/data/LOI14_0000.rad).
DEUXELCT01: two /TH/SHEL (DEF) groups, /TH/PART variables and a /SUBSET
(qa-tests/miniqa/LOIS/LOI66/test_shell/reference, deck test_shell/data/DEUXELC_0000.rad).
Run in the build container, where /root/reference exists.
"""
import os
import shutil

REF = "/root/reference/qa-tests/miniqa/LOIS"
HERE = os.path.dirname(os.path.abspath(__file__))
for src in ("LOI14/reference/LOI14T01", "LOI66/test_shell/reference/DEUXELCT01"):
    shutil.copyfile(os.path.join(REF, src), os.path.join(HERE, os.path.basename(src) + ".th"))
    print("copied", src)
