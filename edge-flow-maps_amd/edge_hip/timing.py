"""Per-launch kernel timing with HIP events on the launch stream.

When enabled, `span(name)` records a start and an end event on the current
stream (the stream every wrapper launches on) around one kernel launch, so
`durations(name)` gives that kernel's launch times in ms without a profiler.
Disabled (the default) it costs nothing.
"""
import contextlib

_enabled = False
_events = {}


def enable(on=True):
    global _enabled
    _enabled = bool(on)


def reset():
    _events.clear()


@contextlib.contextmanager
def span(name):
    if not _enabled:
        yield
        return
    import torch
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    yield
    e1.record()
    _events.setdefault(name, []).append((e0, e1))


def durations(name):
    """ms of every recorded launch of `name` (synchronises on the last event)."""
    evs = _events.get(name, [])
    if evs:
        evs[-1][1].synchronize()
    return [a.elapsed_time(b) for a, b in evs]
