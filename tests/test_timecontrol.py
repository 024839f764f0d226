"""TimeControl schedule (FESolver.cpp:34-56) vs test/TimeControlTest.cpp exact values."""


def test_reference_schedule(orc):
    # TimeControlTest.cpp:9-17: [0, 100], delta 15 -> 15 30 45 60 75 87.5 100, then nextStep is false
    t = orc.timecontrol(0.0, 100.0, 15.0)
    assert list(t) == [15.0, 30.0, 45.0, 60.0, 75.0, 87.5, 100.0]
    assert [x / 100.0 for x in t][:6] == [0.15, 0.30, 0.45, 0.60, 0.75, 0.875]


def test_ten_steps_run_eleven(orc):
    # N=10 on [0,1] takes 11 steps: FP accumulation leaves 0.8999.., then the halving rule
    t = orc.load_steps(10)
    assert len(t) == 11
    assert t[-2] == 0.95 and t[-1] == 1.0


def test_twenty_steps(orc):
    t = orc.load_steps(20)
    assert len(t) == 20 and t[-1] == 1.0
