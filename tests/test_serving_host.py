"""Admission control of the serving loop, host side (no GPU): the memory
gate admits a batch while the running estimates fit the budget (any batch
on an idle device), and the estimator follows measured bytes per unit."""
import threading
import time

from gpu_graph_mining_amd import serving


def test_gate_admits_within_budget_and_waits_for_memory():
    gate = serving.MemoryGate(100)
    stop = threading.Event()
    assert gate.admit_first(80, stop)          # idle device: admitted
    assert not gate.admit_more(30)             # 110 > 100
    assert gate.admit_more(20)                 # 100
    admitted = []

    def second():
        admitted.append(gate.admit_first(50, stop))

    t = threading.Thread(target=second)
    t.start()
    time.sleep(0.2)
    assert not admitted                        # waits: 100 + 50 > 100
    gate.release(100)                          # the first batch ends
    t.join(2)
    assert admitted == [True] and gate.used == 50 and gate.running == 1
    gate.release(50)
    assert gate.admit_first(10_000, stop)      # larger than the budget, but alone
    gate.release(10_000)


def test_gate_wait_ends_when_the_server_stops():
    gate = serving.MemoryGate(10)
    stop = threading.Event()
    assert gate.admit_first(10, stop)
    out = []
    t = threading.Thread(target=lambda: out.append(gate.admit_first(5, stop)))
    t.start()
    time.sleep(0.1)
    stop.set()
    t.join(2)
    assert out == [False]


def test_estimator_follows_measurements():
    e = serving.Estimator(16.0)
    r = serving.Request(serving.QUERY, 0, units=1000)
    assert e.estimate(r) == 16001
    e.observe(1000, 64000)                     # 64 bytes per unit measured
    assert e.per_unit >= 32.0                  # never far below the latest
    for _ in range(30):
        e.observe(1000, 64000)
    assert abs(e.per_unit - 64.0) < 1.0
    e.observe(0, 5)                            # nothing measured: unchanged
    assert abs(e.per_unit - 64.0) < 1.0
