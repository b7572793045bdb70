"""Rewrite the kernel launches of kernels.hip (name<<<grid, block, shm, stream>>>(args);)
into calls of the CPU emulator (emu::launch(grid, block, [=]{ name(args); });).
TEST INFRASTRUCTURE ONLY."""
import re
import sys

src, dst = sys.argv[1], sys.argv[2]
s = open(src).read()
pat = re.compile(r'([A-Za-z_]\w*(?:<[^<>;]*>)?)<<<(.*?)>>>\((.*?)\);', re.S)
s2, n = pat.subn(lambda m: f"emu::launch(emu_cfg({m.group(2)}), [=]() {{ {m.group(1)}({m.group(3)}); }});", s)
s2 = s2.replace('#include "device_types.h"',
                '#include "device_types.h"\n#define emu_cfg(g, b, shm, st) (uint32_t)(g), (uint32_t)(b)', 1)
if n == 0:
    sys.exit("no kernel launches found")
open(dst, "w").write(s2)
