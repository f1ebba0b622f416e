! oracle/ref_units.f90 -- TEST INFRASTRUCTURE ONLY (and the CPU baseline's quiet mode).
! Closes a Fortran unit left open by a reference driver (claw2ez keeps
! fort.info open on unit 10), so its buffered records reach the file.
subroutine ref_close_unit(iu) bind(C, name="ref_close_unit")
  use iso_c_binding, only: c_int
  integer(c_int), value :: iu
  logical :: opened
  inquire(unit=iu, opened=opened)
  if (opened) close(iu)
end subroutine ref_close_unit

! Flushes a unit's buffered records (the reference prints to unit 6).
subroutine ref_flush_unit(iu) bind(C, name="ref_flush_unit")
  use iso_c_binding, only: c_int
  integer(c_int), value :: iu
  flush(iu)
end subroutine ref_flush_unit
