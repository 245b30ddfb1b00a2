"""hipEvent timing on the library's stream (where every kernel runs)."""
from __future__ import annotations

import ctypes as C


class HipEvents:
    def __init__(self, stream: int):
        self.hip = C.CDLL("libamdhip64.so")
        self.stream = C.c_void_p(stream)

    def event(self):
        e = C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e):
        assert self.hip.hipEventRecord(e, self.stream) == 0

    def elapsed_ms(self, a, b) -> float:
        assert self.hip.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), a, b) == 0
        self.hip.hipEventDestroy(a)
        self.hip.hipEventDestroy(b)
        return ms.value

    def device_sync(self):
        assert self.hip.hipDeviceSynchronize() == 0

    def time(self, fn) -> float:
        """Milliseconds between events recorded around fn()."""
        a, b = self.event(), self.event()
        self.record(a)
        fn()
        self.record(b)
        return self.elapsed_ms(a, b)


def time_calls(ev: HipEvents, fn, reps: int) -> list[float]:
    return [ev.time(fn) for _ in range(reps)]
