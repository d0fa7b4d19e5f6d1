! Prints the size of sg_config_t and the byte offsets of selected members,
! which tests/test_fortran_binding.py compares with the C layout.
program fortran_layout_check
  use, intrinsic :: iso_c_binding
  use senga_gpu
  implicit none
  type(sg_config_t), target :: c
  integer(c_intptr_t) :: b
  b = transfer(c_loc(c), b)
  write(*, '(a, i0)') 'size ', c_sizeof(c)
  write(*, '(a, i0)') 'xgdlen ', transfer(c_loc(c%xgdlen), b) - b
  write(*, '(a, i0)') 'bc_rpar ', transfer(c_loc(c%bc_rpar), b) - b
  write(*, '(a, i0)') 'yrin ', transfer(c_loc(c%yrin), b) - b
  write(*, '(a, i0)') 'ncpoly ', transfer(c_loc(c%ncpoly), b) - b
  write(*, '(a, i0)') 'amascp ', transfer(c_loc(c%amascp), b) - b
  write(*, '(a, i0)') 'rparam ', transfer(c_loc(c%rparam), b) - b
  write(*, '(a, i0)') 'mslist ', transfer(c_loc(c%mslist), b) - b
  write(*, '(a, i0)') 'effy3b ', transfer(c_loc(c%effy3b), b) - b
  write(*, '(a, i0)') 'clnten ', transfer(c_loc(c%clnten), b) - b
  write(*, '(a, i0)') 'ncovis ', transfer(c_loc(c%ncovis), b) - b
  write(*, '(a, i0)') 'diffco ', transfer(c_loc(c%diffco), b) - b
  write(*, '(a, i0)') 'wilko2 ', transfer(c_loc(c%wilko2), b) - b
end program fortran_layout_check
