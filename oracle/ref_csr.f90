!----------------------------------------------------------------------
! CSR wake harness for the IMPACT-Z reference (test infrastructure only).
!
! Calls the reference's own csrwakeTrIGF_FieldQuant (src/Appl/Field.f90:
! 1007-1251) on the line densities of a case file and writes the wakes.
!
! Input (list-directed, file named on the command line): ncase, then per
! case:  Nx r0 ptmin hx blength gam  followed by Nx line densities.
! Output (stream, file named second): per case Nx doubles of ezwake.
!----------------------------------------------------------------------
      program ref_csr
        use FieldQuantclass
        implicit none
        integer :: ncase, icase, nx, u
        double precision :: r0, ptmin, hx, blength, gam
        double precision, dimension(:), allocatable :: rho, rhop, rhopp, ez
        character(len=256) :: fin, fout

        call get_command_argument(1, fin)
        call get_command_argument(2, fout)
        open(unit=11, file=trim(fin), status='old')
        open(newunit=u, file=trim(fout), access='stream', form='unformatted', &
             status='replace')
        read(11,*) ncase
        do icase = 1, ncase
          read(11,*) nx, r0, ptmin, hx, blength, gam
          allocate(rho(nx), rhop(nx), rhopp(nx), ez(nx))
          read(11,*) rho
          rhop = 0.0d0
          rhopp = 0.0d0
          ez = 0.0d0
          call csrwakeTrIGF_FieldQuant(nx, r0, ptmin, hx, blength, rho, rhop, rhopp, &
                                       gam, ez)
          write(u) ez
          deallocate(rho, rhop, rhopp, ez)
        enddo
        close(u)
        close(11)
      end program ref_csr
