# Restart snapshots for a Julia host, in the HDF5 file this package writes
# (hdf5io.save_solution_hdf5 / GodunovSolver.save_snapshot): the fields of the
# reference's solutionCellsT (primeObjects.jl:197-205) as HDF5 datasets dt,
# flowTime, nCells, densityCells, UxCells, UyCells, pressureCells.  The
# reference writes that struct with BSON.jl's @save (evaluate2d.jl:187-203)
# and reads it back in loadPrevResults (loadPrevResults.jl:3-38); these two
# functions are that pair over HDF5.jl, which the reference already uses for
# its meshes (utilsIO.jl:107-316).
module LibGodunovIO

using HDF5

export saveSolutionHDF5, loadPrevResultsHDF5

"saveSolutionHDF5(filename, dt, flowTime, testfields2d): the cell primitives and times."
function saveSolutionHDF5(filename::String, dt::Float64, flowTime::Float64, f)
    h5open(filename, "w") do io
        write(io, "dt", dt)
        write(io, "flowTime", flowTime)
        write(io, "nCells", Int64(length(f.densityCells)))
        write(io, "densityCells", f.densityCells)
        write(io, "UxCells", f.UxCells)
        write(io, "UyCells", f.UyCells)
        write(io, "pressureCells", f.pressureCells)
    end
    nothing
end

"""loadPrevResultsHDF5(testMesh, thermo, filename, dynControls, testfields2dX):
loadPrevResults (loadPrevResults.jl:3-38) from a snapshot of this package:
the cell primitives, aSoundCells and VMAXCells as the reference derives them
(loadPrevResults.jl:18-19), and dynControls.flowTime."""
function loadPrevResultsHDF5(testMesh, thermo, filename::String, dynControls, testfields2dX)
    h5open(filename, "r") do io
        n = read(io, "nCells")
        n == testMesh.nCells || error("$filename holds $n cells, the mesh $(testMesh.nCells)")
        f = testfields2dX
        f.densityCells .= read(io, "densityCells")
        f.UxCells .= read(io, "UxCells")
        f.UyCells .= read(io, "UyCells")
        f.pressureCells .= read(io, "pressureCells")
        for i in 1:n
            f.aSoundCells[i] = sqrt(thermo.Gamma * f.pressureCells[i] / f.densityCells[i])
            f.VMAXCells[i] = sqrt(f.UxCells[i] * f.UxCells[i] + f.UyCells[i] * f.UyCells[i]) + f.aSoundCells[i]
        end
        dynControls.flowTime = read(io, "flowTime")
    end
    nothing
end

end  # module
