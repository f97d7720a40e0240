"""Device buffers passed to a C entry stay alive through the call (the
regression test of a launch whose inline temporaries were freed before the
kernel read them)."""

import gc
import weakref

from felizjueves_amd import jk


class FakeTensor:
    def __init__(self, addr):
        self.addr = addr

    def data_ptr(self):
        return self.addr


def test_marshal_keeps_tensors_referenced():
    made = []

    def make(addr):
        t = FakeTensor(addr)
        made.append(weakref.ref(t))
        return t

    args = (make(0x1000), 7, make(0x2000))
    vals = jk._marshal(args)
    gc.collect()
    # while `args` is held (as _call holds it through the C call) nothing
    # it references has been freed
    assert all(r() is not None for r in made)
    assert vals[0].value == 0x1000 and vals[2].value == 0x2000 and vals[1] == 7
    del args
    gc.collect()
    assert all(r() is None for r in made)
