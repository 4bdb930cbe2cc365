"""CPU oracle of the reference MLS-MPM substep. Test infrastructure only (tests, smoke check, CPU baseline)."""
